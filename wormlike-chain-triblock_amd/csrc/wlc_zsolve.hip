// Kernel B of the propagator step for one Lmax (this file is compiled once per L
// with -DWLC_L=<L>): z pass of the forward FFT, k-aligned solve of G(k) q = P(k)
// (replaces the per-k dense G^-1 apply of wormlike_MDE_2dcomp_mod.f90:395-409 with
// the tables built by init_inverse_sparse_GAB, Gmatrix_2dcomp_mod.f90:222-650),
// inverse z pass.
//
// One thread owns one k-point (kx, ky, kz) and holds all M harmonics of it in
// registers.  The half spectrum is laid out [i][ky][kx/4][z][kx%4]: a tile of four
// kx columns of one harmonic is contiguous (4 nz complex values).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "wlc_device.h"

#ifndef WLC_L
#error "compile with -DWLC_L=<Lmax>"
#endif

#define WLC_CAT2(a, b) a##b
#define WLC_CAT(a, b) WLC_CAT2(a, b)

namespace wlc {
namespace WLC_CAT(per_l_, WLC_L) {

namespace {
constexpr int kL = WLC_L;
constexpr int kM = (kL + 1) * (kL + 1);
constexpr int kHB = 8;  // harmonics per LDS batch of the z transform
}  // namespace

template <int DIR>
__device__ __forceinline__ void z_transform(double2 (&v)[kM], double2* lds, const AxisDev& az,
                                            int TX, int lidx) {
  const int nz = az.n;
#pragma unroll
  for (int b0 = 0; b0 < kM; b0 += kHB) {
    const int nb = (kM - b0) < kHB ? (kM - b0) : kHB;
#pragma unroll
    for (int h = 0; h < kHB; ++h)
      if (b0 + h < kM) lds[h * TX * nz + lidx] = v[b0 + h];
    __syncthreads();
    fft_axis<DIR, kHB>(lds, az, nb * TX, 1, nz);
#pragma unroll
    for (int h = 0; h < kHB; ++h)
      if (b0 + h < kM) v[b0 + h] = lds[h * TX * nz + lidx];
    __syncthreads();
  }
}

// generic kernel B: any nz <= 256; TX columns of one ky row per workgroup, z transforms
// through LDS eight harmonics at a time (block-wide Stockham passes)
template <class S>
__device__ __forceinline__ void k_zsolve_body(const Geom& g, const SolveArgs& a) {
  extern __shared__ double2 lds[];
  const int nz = g.nz, ny = g.ny, nxh = g.nxh;
  const int TX = a.TX;  // TX * nz == blockDim.x
  const int t = threadIdx.x;
  const int tx = t / nz, z = t - tx * nz;
  const int kx = blockIdx.x * TX + tx, ky = blockIdx.y;
  const bool valid = kx < nxh;
  const long long istride = spec_index(1, 0, 0, 0, ny, g.nt, nz);
  S* p = reinterpret_cast<S*>(a.spec) + spec_index(0, ky, valid ? kx : 0, z, ny, g.nt, nz);
  double2 v[kM];
#pragma unroll
  for (int i = 0; i < kM; ++i) v[i] = valid ? spec_ld(p + i * istride) : make_double2(0.0, 0.0);
  const int lidx = tx * nz + z;
  z_transform<-1>(v, lds, g.az, TX, lidx);
  if (a.solve && valid) {
    const double kxv = signed_k(kx, g.nx) * a.kfac[0];
    const double kyv = signed_k(ky, ny) * a.kfac[1];
    const double kzv = signed_k(z, nz) * a.kfac[2];
    Solver<kL>::solve(v, kxv, kyv, kzv, a.cdelta, a.sigma, a.ds, a.dslam);
  }
  z_transform<+1>(v, lds, g.az, TX, lidx);
  if (valid) {
#pragma unroll
    for (int i = 0; i < kM; ++i) spec_st(p + i * istride, v[i]);
  }
}

template <class S>
__global__ void __launch_bounds__(256) k_zsolve(Geom g, SolveArgs a) {
  k_zsolve_body<S>(g, a);
}

// the same for one instance of a batch per blockIdx.z
template <class S>
__global__ void __launch_bounds__(256) k_zsolve_b(const ZJob* __restrict__ jobs) {
  const ZJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_zsolve_body<S>(j.g, j.a);
}

// kernel B for nz = 64: one wave owns one (kx, ky) column, lane = z; the 64-point z
// transform is an 8 x 8 four-step done wave-locally in LDS, eight harmonics per
// batch; each lane's step-1 twiddles stay in registers.  LDS rows of 8 are padded
// to 9 so the strided step-2 reads are bank-conflict free.
__device__ __forceinline__ int pad8(int u) { return u + (u >> 3); }
constexpr int kPadBatch = 72;  // padded units per harmonic (64 + 8)

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// wave_sync that also keeps the instruction scheduler from moving work across it (SB):
// with all 81 harmonics of Lmax 8 in registers, scheduling freedom across the transform
// steps costs spills
template <bool SB>
__device__ __forceinline__ void wave_sync_sb() {
  wave_sync();
  if constexpr (SB) __builtin_amdgcn_sched_barrier(0);
}

// Workgroup = one kx tile (4 columns) of one ky row; wave w owns column kx = 4 t + w.
// Per batch of 8 harmonics the four waves cooperatively load each harmonic's 4 KB tile
// (64 z x 4 kx, contiguous) with 1 KB-coalesced loads into LDS, wave w's region holding
// its column in natural z order; the next batch is prefetched into registers meanwhile.
// The inverse writes the step-2 outputs back into the regions and the workgroup stores
// the tiles with coalesced 1 KB stores.
constexpr int kRegion = kHB * kPadBatch + 2;  // per-wave LDS region (padded stride)

template <class S>
__device__ __forceinline__ void tile_load(double2 (&r)[kHB], const S* __restrict__ tile,
                                          long long istride, int b0, int t) {
#pragma unroll
  for (int h = 0; h < kHB; ++h)
    r[h] = (b0 + h < kM) ? spec_ld(tile + (long long)(b0 + h) * istride + t) : make_double2(0.0, 0.0);
}

__device__ __forceinline__ void tile_to_lds(const double2 (&r)[kHB], double2* lds, int t) {
  const int z = t >> 2, w = t & 3;
#pragma unroll
  for (int h = 0; h < kHB; ++h) lds[w * kRegion + h * kPadBatch + pad8(z)] = r[h];
}

template <class S>
__device__ __forceinline__ void z64_forward(double2 (&v)[kM], double2* lds, int wv, int lane,
                                            const double2 (&tw)[8], const S* __restrict__ tile,
                                            long long istride) {
  const int h = lane >> 3, j = lane & 7, t = threadIdx.x;
  double2* wl = lds + wv * kRegion;
  double2 nxt[kHB];
  tile_load(nxt, tile, istride, 0, t);
#pragma unroll
  for (int b0 = 0; b0 < kM; b0 += kHB) {
    __syncthreads();  // every wave is done with the previous batch's regions
    tile_to_lds(nxt, lds, t);
    if (b0 + kHB < kM) tile_load(nxt, tile, istride, b0 + kHB, t);
    __syncthreads();
    double2 u[8];
#pragma unroll
    for (int n2 = 0; n2 < 8; ++n2) u[n2] = wl[h * kPadBatch + pad8(j + 8 * n2)];
    Dft<8, -1>::run(u);
#pragma unroll
    for (int k2 = 1; k2 < 8; ++k2) u[k2] = twiddle<-1>(u[k2], tw[k2]);
    wave_sync();
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) wl[h * kPadBatch + pad8(8 * k2 + j)] = u[k2];
    wave_sync();
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) u[jj] = wl[h * kPadBatch + pad8(8 * j + jj)];
    Dft<8, -1>::run(u);
    wave_sync();
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) wl[h * kPadBatch + pad8(j + 8 * k1)] = u[k1];
    wave_sync();
#pragma unroll
    for (int q = 0; q < kHB; ++q)
      if (b0 + q < kM) v[b0 + q] = wl[q * kPadBatch + pad8(lane)];
  }
}

template <class S>
__device__ __forceinline__ void z64_inverse(const double2 (&v)[kM], double2* lds, int wv, int lane,
                                            const double2 (&tw)[8], S* __restrict__ tile,
                                            long long istride) {
  const int h = lane >> 3, j = lane & 7, t = threadIdx.x;
  const int tz = t >> 2, tw4 = t & 3;
  double2* wl = lds + wv * kRegion;
#pragma unroll
  for (int b0 = 0; b0 < kM; b0 += kHB) {
    __syncthreads();  // the previous batch's tiles have been read out of LDS
#pragma unroll
    for (int q = 0; q < kHB; ++q)
      if (b0 + q < kM) wl[q * kPadBatch + pad8(lane)] = v[b0 + q];
    wave_sync();
    double2 u[8];
#pragma unroll
    for (int n2 = 0; n2 < 8; ++n2) u[n2] = wl[h * kPadBatch + pad8(j + 8 * n2)];
    Dft<8, +1>::run(u);
#pragma unroll
    for (int k2 = 1; k2 < 8; ++k2) u[k2] = twiddle<+1>(u[k2], tw[k2]);
    wave_sync();
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) wl[h * kPadBatch + pad8(8 * k2 + j)] = u[k2];
    wave_sync();
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) u[jj] = wl[h * kPadBatch + pad8(8 * j + jj)];
    Dft<8, +1>::run(u);
    wave_sync();
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) wl[h * kPadBatch + pad8(j + 8 * k1)] = u[k1];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kHB; ++q)
      if (b0 + q < kM)
        spec_st(tile + (long long)(b0 + q) * istride + t, lds[tw4 * kRegion + q * kPadBatch + pad8(tz)]);
  }
}

template <int OCC, class S>
__device__ __forceinline__ void k_zsolve64_body(const Geom& g, const SolveArgs& a) {
  __shared__ double2 lds[4 * kRegion];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ny = g.ny, nxh = g.nxh;
  const int t4 = blockIdx.x, ky = blockIdx.y;
  const int kx = t4 * kKxTile + wv;
  const bool valid = kx < nxh;
  const long long istride = spec_index(1, 0, 0, 0, ny, g.nt, 64);
  // the whole 4-kx tile is allocated even where kx >= nxh (the layout pads to 4)
  S* tile = reinterpret_cast<S*>(a.spec) + spec_index(0, ky, t4 * kKxTile, 0, ny, g.nt, 64);
  double2 tw[8];
  {
    const int j = lane & 7;
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) tw[k2] = g.az.tw[j * k2];
  }
  double2 v[kM];
  z64_forward(v, lds, wv, lane, tw, tile, istride);
  if (a.solve && valid) {
    const double kxv = signed_k(kx, g.nx) * a.kfac[0];
    const double kyv = signed_k(ky, ny) * a.kfac[1];
    const double kzv = signed_k(lane, 64) * a.kfac[2];
    Solver<kL>::solve(v, kxv, kyv, kzv, a.cdelta, a.sigma, a.ds, a.dslam);
  }
  z64_inverse(v, lds, wv, lane, tw, tile, istride);
}

template <int OCC, class S>
__global__ void __launch_bounds__(256, OCC) k_zsolve64(Geom g, SolveArgs a) {
  k_zsolve64_body<OCC, S>(g, a);
}

// the same for one instance of a batch per blockIdx.z
template <int OCC, class S>
__global__ void __launch_bounds__(256, OCC) k_zsolve64_b(const ZJob* __restrict__ jobs) {
  const ZJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_zsolve64_body<OCC, S>(j.g, j.a);
}

// ---------------------------------------------------------------------------
// nz = 64, 5 <= Lmax <= 7: the harmonics of a column split over two waves by degree
// ---------------------------------------------------------------------------
// Lo holds l = 0..kLS, Hi holds l = kLS+1..L; each wave keeps only its part in registers,
// which allows two waves per SIMD instead of one (at Lmax = 8 the larger part alone does
// not fit two waves per SIMD, so Lmax = 8 keeps the one-wave kernel).  Workgroup = two kx columns of a
// 4-kx tile (waves: column c = w >> 1, part = w & 1).  The two workgroups of a tile run
// back to back on one XCD (block order below), so each tile's cache lines are fetched once.
constexpr int kLS = kL >= 7 ? 5 : (kL == 6 ? 4 : 3);
using PartLo = SolverPart<kL, 0, kLS>;
using PartHi = SolverPart<kL, kLS + 1, kL>;
constexpr int kNV = PartLo::NH > PartHi::NH ? PartLo::NH : PartHi::NH;
constexpr int kXch = (2 * kLS + 1) * 64;                       // crossing chains x lanes
constexpr int kRegionS = kRegion > kXch + 2 ? kRegion : kXch + 2;  // per-wave LDS region

// forward z transform of the NH harmonics in v (lane = z -> lane = kz), batches of 8
template <int NH, bool SB = false, int N>
__device__ __forceinline__ void z64_fwd_regs(double2 (&v)[N], double2* wl, int lane,
                                             const double2 (&tw)[8]) {
  const int h = lane >> 3, j = lane & 7;
#pragma unroll
  for (int b0 = 0; b0 < NH; b0 += kHB) {
#pragma unroll
    for (int q = 0; q < kHB; ++q)
      if (b0 + q < NH) wl[q * kPadBatch + pad8(lane)] = v[b0 + q];
    wave_sync_sb<SB>();
    double2 u[8];
#pragma unroll
    for (int n2 = 0; n2 < 8; ++n2) u[n2] = wl[h * kPadBatch + pad8(j + 8 * n2)];
    Dft<8, -1>::run(u);
#pragma unroll
    for (int k2 = 1; k2 < 8; ++k2) u[k2] = twiddle<-1>(u[k2], tw[k2]);
    wave_sync_sb<SB>();
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) wl[h * kPadBatch + pad8(8 * k2 + j)] = u[k2];
    wave_sync_sb<SB>();
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) u[jj] = wl[h * kPadBatch + pad8(8 * j + jj)];
    Dft<8, -1>::run(u);
    wave_sync_sb<SB>();
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) wl[h * kPadBatch + pad8(j + 8 * k1)] = u[k1];
    wave_sync_sb<SB>();
#pragma unroll
    for (int q = 0; q < kHB; ++q)
      if (b0 + q < NH) v[b0 + q] = wl[q * kPadBatch + pad8(lane)];
    wave_sync_sb<SB>();
  }
}

// inverse z transform (lane = kz -> lane = z), batches of 8
template <int NH, bool SB = false, int N>
__device__ __forceinline__ void z64_inv_regs(double2 (&v)[N], double2* wl, int lane,
                                             const double2 (&tw)[8]) {
  const int h = lane >> 3, j = lane & 7;
#pragma unroll
  for (int b0 = 0; b0 < NH; b0 += kHB) {
#pragma unroll
    for (int q = 0; q < kHB; ++q)
      if (b0 + q < NH) wl[q * kPadBatch + pad8(lane)] = v[b0 + q];
    wave_sync_sb<SB>();
    double2 u[8];
#pragma unroll
    for (int n2 = 0; n2 < 8; ++n2) u[n2] = wl[h * kPadBatch + pad8(j + 8 * n2)];
    Dft<8, +1>::run(u);
#pragma unroll
    for (int k2 = 1; k2 < 8; ++k2) u[k2] = twiddle<+1>(u[k2], tw[k2]);
    wave_sync_sb<SB>();
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) wl[h * kPadBatch + pad8(8 * k2 + j)] = u[k2];
    wave_sync_sb<SB>();
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) u[jj] = wl[h * kPadBatch + pad8(8 * j + jj)];
    Dft<8, +1>::run(u);
    wave_sync_sb<SB>();
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) wl[h * kPadBatch + pad8(j + 8 * k1)] = u[k1];
    wave_sync_sb<SB>();
#pragma unroll
    for (int q = 0; q < kHB; ++q)
      if (b0 + q < NH) v[b0 + q] = wl[q * kPadBatch + pad8(lane)];
    wave_sync_sb<SB>();
  }
}

template <int OCC, class S>
__device__ __forceinline__ void k_zsolve64s_body(const Geom& g, const SolveArgs& a) {
  __shared__ double2 lds[4 * kRegionS];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int c = wv >> 1, part = wv & 1;
  const int ny = g.ny, nxh = g.nxh, nt = g.nt;
  // block -> (tile, half): blocks b and b + 8 (consecutive on XCD b mod 8) share a tile
  const int nb = gridDim.x, b = blockIdx.x;
  int tl, half;
  if ((nb & 15) == 0) {
    const int t = b >> 3;
    half = t & 1;
    tl = (t >> 1) * 8 + (b & 7);
  } else {
    half = b & 1;
    tl = b >> 1;
  }
  const int t4 = tl % nt, ky = tl / nt;
  // both columns of this workgroup lie in the padding of the last tile (kx >= nx/2 + 1):
  // nothing to do (the condition is uniform over the workgroup, before any barrier)
  if (t4 * kKxTile + 2 * half >= nxh) return;
  const int kxl = 2 * half + c;
  const int kx = t4 * kKxTile + kxl;
  const bool real_col = kx < nxh;  // wave-uniform; padding columns are neither read nor written
  const bool active = a.solve && real_col;
  const long long istride = spec_index(1, 0, 0, 0, ny, nt, 64);
  // column kx of the tile: element (i, z) at i * istride + 4 z (the tile is allocated
  // even where kx >= nxh)
  S* col = reinterpret_cast<S*>(a.spec) + spec_index(0, ky, t4 * kKxTile, 0, ny, nt, 64) + kxl;
  double2* wl = lds + wv * kRegionS;
  double2* xch = lds + (wv & ~1) * kRegionS;  // the Lo wave's region of this column
  double2 tw[8];
  {
    const int j = lane & 7;
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) tw[k2] = g.az.tw[j * k2];
  }
  const KFrame f(signed_k(kx, g.nx) * a.kfac[0], signed_k(ky, ny) * a.kfac[1],
                 signed_k(lane, 64) * a.kfac[2], a.sigma, a.ds);
  double2 v[kNV];
  // phase 1: load, forward z transform, rotate into the k frame; Lo: forward elimination
  if (part == 0 && real_col) {
#pragma unroll
    for (int h = 0; h < PartLo::NH; ++h) v[h] = spec_ld(col + (long long)(PartLo::B0 + h) * istride + 4 * lane);
    z64_fwd_regs<PartLo::NH>(v, wl, lane, tw);
    if (active) {
      PartLo::to_frame(v, f);
      tri_lower_forward<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
    }
  } else if (real_col) {
#pragma unroll
    for (int h = 0; h < PartHi::NH; ++h) v[h] = spec_ld(col + (long long)(PartHi::B0 + h) * istride + 4 * lane);
    z64_fwd_regs<PartHi::NH>(v, wl, lane, tw);
    if (active) PartHi::to_frame(v, f);
  }
  __syncthreads();
  // phase 2: Hi finishes every chain
  if (part == 1 && active) tri_upper<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
  __syncthreads();
  // phase 3: Lo back substitution; rotate back, inverse z transform, store
  if (part == 0 && real_col) {
    if (active) {
      tri_lower_back<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
      PartLo::from_frame(v, f);
    }
    z64_inv_regs<PartLo::NH>(v, wl, lane, tw);
#pragma unroll
    for (int h = 0; h < PartLo::NH; ++h) spec_st(col + (long long)(PartLo::B0 + h) * istride + 4 * lane, v[h]);
  } else if (real_col) {
    if (active) PartHi::from_frame(v, f);
    z64_inv_regs<PartHi::NH>(v, wl, lane, tw);
#pragma unroll
    for (int h = 0; h < PartHi::NH; ++h) spec_st(col + (long long)(PartHi::B0 + h) * istride + 4 * lane, v[h]);
  }
}

template <int OCC, class S>
__global__ void __launch_bounds__(256, OCC) k_zsolve64s(Geom g, SolveArgs a) {
  k_zsolve64s_body<OCC, S>(g, a);
}

// the same for one instance of a batch per blockIdx.z
template <int OCC, class S>
__global__ void __launch_bounds__(256, OCC) k_zsolve64s_b(const ZJob* __restrict__ jobs) {
  const ZJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_zsolve64s_body<OCC, S>(j.g, j.a);
}

// ---------------------------------------------------------------------------
// nz = 64, 5 <= Lmax <= 7: one 4-kx tile per workgroup, harmonics split by degree
// ---------------------------------------------------------------------------
// Eight waves: wave w = 4 p + c holds part p (0: l <= kLS, 1: l > kLS) of column
// kx = 4 t + c.  The four waves of a part move each of their harmonics' 4 KB tiles
// (64 z x 4 kx, contiguous) between HBM and their four LDS regions with coalesced
// 16-byte accesses (thread i of the part: z = i >> 2, column i & 3), eight harmonics per
// batch, the next batch's loads in flight while the current batch is transformed.  Both
// parts run kNBt batches (the smaller part idles through its last ones), so every wave
// passes the same barriers.  Columns beyond nx/2 (the padding of a partial last tile)
// are neither read, transformed nor written.
constexpr int kNBt = (kNV + kHB - 1) / kHB;
constexpr int kTileLds = 2 * kKxTile * kRegionS;  // dynamic LDS, double2 units

// NCOL columns per part (4: the whole tile, k_zsolve64t; 2: half a tile, k_zsolve64p):
// thread gi of the part moves element (z = gi / NCOL, column gi % NCOL) of each harmonic, at
// offset 4 z + column of the part's first column (hp)
template <int NCOL = kKxTile, class S>
__device__ __forceinline__ void zt_forward(double2 (&v)[kNV], double2* glds, double2* wl, int lane,
                                           int gi, int nh, bool run, const double2 (&tw)[8],
                                           const S* __restrict__ hp, long long istride, bool io) {
  const int h = lane >> 3, j = lane & 7;
  const int tz = gi / NCOL, tc = gi % NCOL;
  const int go = kKxTile * tz + tc;  // the element's offset in a harmonic's tile
  double2 nxt[kHB];
#pragma unroll
  for (int q = 0; q < kHB; ++q)
    nxt[q] = (io && q < nh) ? spec_ld(hp + (long long)q * istride + go) : make_double2(0.0, 0.0);
#pragma unroll
  for (int bb = 0; bb < kNBt; ++bb) {
    const int b0 = bb * kHB;
    __syncthreads();  // every wave of the part has read the previous batch out of its region
#pragma unroll
    for (int q = 0; q < kHB; ++q)
      if (b0 + q < nh) glds[tc * kRegionS + q * kPadBatch + pad8(tz)] = nxt[q];
    if (bb + 1 < kNBt) {
#pragma unroll
      for (int q = 0; q < kHB; ++q) {
        const int hq = b0 + kHB + q;
        nxt[q] = (io && hq < nh) ? spec_ld(hp + (long long)hq * istride + go) : make_double2(0.0, 0.0);
      }
    }
    __syncthreads();
    if (run && b0 < nh) {
      double2 u[8];
#pragma unroll
      for (int n2 = 0; n2 < 8; ++n2) u[n2] = wl[h * kPadBatch + pad8(j + 8 * n2)];
      Dft<8, -1>::run(u);
#pragma unroll
      for (int k2 = 1; k2 < 8; ++k2) u[k2] = twiddle<-1>(u[k2], tw[k2]);
      wave_sync();
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) wl[h * kPadBatch + pad8(8 * k2 + j)] = u[k2];
      wave_sync();
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) u[jj] = wl[h * kPadBatch + pad8(8 * j + jj)];
      Dft<8, -1>::run(u);
      wave_sync();
#pragma unroll
      for (int k1 = 0; k1 < 8; ++k1) wl[h * kPadBatch + pad8(j + 8 * k1)] = u[k1];
      wave_sync();
#pragma unroll
      for (int q = 0; q < kHB; ++q)
        if (b0 + q < kNV && b0 + q < nh) v[b0 + q] = wl[q * kPadBatch + pad8(lane)];
    }
  }
}

template <int NCOL = kKxTile, class S>
__device__ __forceinline__ void zt_inverse(const double2 (&v)[kNV], double2* glds, double2* wl,
                                           int lane, int gi, int nh, bool run,
                                           const double2 (&tw)[8], S* __restrict__ hp,
                                           long long istride, bool io) {
  const int h = lane >> 3, j = lane & 7;
  const int tz = gi / NCOL, tc = gi % NCOL;
  const int go = kKxTile * tz + tc;
#pragma unroll
  for (int bb = 0; bb < kNBt; ++bb) {
    const int b0 = bb * kHB;
    __syncthreads();  // the previous batch's tiles have been read out of the regions
    if (run && b0 < nh) {
#pragma unroll
      for (int q = 0; q < kHB; ++q)
        if (b0 + q < kNV && b0 + q < nh) wl[q * kPadBatch + pad8(lane)] = v[b0 + q];
      wave_sync();
      double2 u[8];
#pragma unroll
      for (int n2 = 0; n2 < 8; ++n2) u[n2] = wl[h * kPadBatch + pad8(j + 8 * n2)];
      Dft<8, +1>::run(u);
#pragma unroll
      for (int k2 = 1; k2 < 8; ++k2) u[k2] = twiddle<+1>(u[k2], tw[k2]);
      wave_sync();
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) wl[h * kPadBatch + pad8(8 * k2 + j)] = u[k2];
      wave_sync();
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) u[jj] = wl[h * kPadBatch + pad8(8 * j + jj)];
      Dft<8, +1>::run(u);
      wave_sync();
#pragma unroll
      for (int k1 = 0; k1 < 8; ++k1) wl[h * kPadBatch + pad8(j + 8 * k1)] = u[k1];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kHB; ++q)
      if (io && b0 + q < nh)
        spec_st(hp + (long long)(b0 + q) * istride + go, glds[tc * kRegionS + q * kPadBatch + pad8(tz)]);
  }
}

template <class S>
__device__ __forceinline__ void k_zsolve64t_body(const Geom& g, const SolveArgs& a) {
  extern __shared__ double2 lds[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int part = wv >> 2, c = wv & 3, gi = threadIdx.x & 255;
  const int ny = g.ny, nxh = g.nxh, nt = g.nt;
  // tiles with four real columns first; the partial last tile of every ky row (when
  // nx/2 + 1 is not a multiple of 4) last, so the final, partly filled round of
  // workgroups holds the cheap ones
  const int nfull = nxh / kKxTile;
  const int b = blockIdx.x;
  int t4, ky;
  if (b < nfull * ny) {
    t4 = b % nfull;
    ky = b / nfull;
  } else {
    t4 = nfull;
    ky = b - nfull * ny;
  }
  const int kx = t4 * kKxTile + c;
  const bool real_col = kx < nxh;                       // wave-uniform
  const bool io = t4 * kKxTile + (gi & 3) < nxh;        // the column this thread moves
  const bool active = a.solve && real_col;
  const long long istride = spec_index(1, 0, 0, 0, ny, nt, 64);
  const int nh = part ? PartHi::NH : PartLo::NH;
  const int hb0 = part ? PartHi::B0 : PartLo::B0;
  S* hp = reinterpret_cast<S*>(a.spec) + spec_index(0, ky, t4 * kKxTile, 0, ny, nt, 64) +
          (long long)hb0 * istride;
  double2* glds = lds + part * (kKxTile * kRegionS);
  double2* wl = glds + c * kRegionS;
  double2* xch = lds + c * kRegionS;  // the Lo wave's region of column c
  double2 tw[8];
  {
    const int jt = lane & 7;
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) tw[k2] = g.az.tw[jt * k2];
  }
  const KFrame f(signed_k(kx, g.nx) * a.kfac[0], signed_k(ky, ny) * a.kfac[1],
                 signed_k(lane, 64) * a.kfac[2], a.sigma, a.ds);
  double2 v[kNV];
  // phase 1: load, forward z transform, rotate into the k frame; Lo: forward elimination
  zt_forward(v, glds, wl, lane, gi, nh, real_col, tw, hp, istride, io);
  if (active) {
    if (part == 0) {
      PartLo::to_frame(v, f);
      tri_lower_forward<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
    } else {
      PartHi::to_frame(v, f);
    }
  }
  __syncthreads();
  // phase 2: Hi finishes every chain
  if (part == 1 && active) tri_upper<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
  __syncthreads();
  // phase 3: Lo back substitution; rotate back, inverse z transform, store
  if (active) {
    if (part == 0) {
      tri_lower_back<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
      PartLo::from_frame(v, f);
    } else {
      PartHi::from_frame(v, f);
    }
  }
  zt_inverse(v, glds, wl, lane, gi, nh, real_col, tw, hp, istride, io);
}

template <class S>
__global__ void __launch_bounds__(512) k_zsolve64t(Geom g, SolveArgs a) {
  k_zsolve64t_body<S>(g, a);
}

// the same for one instance of a batch per blockIdx.z
template <class S>
__global__ void __launch_bounds__(512) k_zsolve64t_b(const ZJob* __restrict__ jobs) {
  const ZJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_zsolve64t_body<S>(j.g, j.a);
}

// Half-tile variant (k_zsolve64p): the same split, coalesced through LDS, with a workgroup
// of four waves per two adjacent columns of a tile (wave w = 2 p + c: part p of column
// 2 half + c), so that two workgroups fit on a CU (217 VGPRs, two waves per SIMD) and one
// workgroup's loads and stores overlap the other's transforms and solve; the whole-tile
// kernel has a single eight-wave workgroup per CU, whose phases cannot overlap.  Each
// part moves 32 contiguous bytes per z of every harmonic; the two halves of a tile are
// blocks b and b + 8, consecutive on one XCD (the same L2).
template <class S>
__device__ __forceinline__ void k_zsolve64p_body(const Geom& g, const SolveArgs& a) {
  extern __shared__ double2 lds[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int part = wv >> 1, c = wv & 1, gi = threadIdx.x & 127;
  const int ny = g.ny, nxh = g.nxh, nt = g.nt;
  const int nb = gridDim.x, b = blockIdx.x;
  int tl, half;
  if ((nb & 15) == 0) {
    const int t = b >> 3;
    half = t & 1;
    tl = (t >> 1) * 8 + (b & 7);
  } else {
    half = b & 1;
    tl = b >> 1;
  }
  const int t4 = tl % nt, ky = tl / nt;
  if (t4 * kKxTile + 2 * half >= nxh) return;  // both columns are padding (uniform)
  const int kx = t4 * kKxTile + 2 * half + c;
  const bool real_col = kx < nxh;                                       // wave-uniform
  const bool io = t4 * kKxTile + 2 * half + (gi & 1) < nxh;             // the column this thread moves
  const bool active = a.solve && real_col;
  const long long istride = spec_index(1, 0, 0, 0, ny, nt, 64);
  const int nh = part ? PartHi::NH : PartLo::NH;
  const int hb0 = part ? PartHi::B0 : PartLo::B0;
  S* hp = reinterpret_cast<S*>(a.spec) + spec_index(0, ky, t4 * kKxTile, 0, ny, nt, 64) + 2 * half +
          (long long)hb0 * istride;
  double2* glds = lds + part * (2 * kRegionS);
  double2* wl = glds + c * kRegionS;
  double2* xch = lds + c * kRegionS;  // the Lo wave's region of column c
  double2 tw[8];
  {
    const int jt = lane & 7;
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) tw[k2] = g.az.tw[jt * k2];
  }
  const KFrame f(signed_k(kx, g.nx) * a.kfac[0], signed_k(ky, ny) * a.kfac[1],
                 signed_k(lane, 64) * a.kfac[2], a.sigma, a.ds);
  double2 v[kNV];
  zt_forward<2>(v, glds, wl, lane, gi, nh, real_col, tw, hp, istride, io);
  if (active) {
    if (part == 0) {
      PartLo::to_frame(v, f);
      tri_lower_forward<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
    } else {
      PartHi::to_frame(v, f);
    }
  }
  __syncthreads();
  if (part == 1 && active) tri_upper<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
  __syncthreads();
  if (active) {
    if (part == 0) {
      tri_lower_back<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
      PartLo::from_frame(v, f);
    } else {
      PartHi::from_frame(v, f);
    }
  }
  zt_inverse<2>(v, glds, wl, lane, gi, nh, real_col, tw, hp, istride, io);
}

template <class S>
__global__ void __launch_bounds__(256, 2) k_zsolve64p(Geom g, SolveArgs a) {
  k_zsolve64p_body<S>(g, a);
}

// the same for one instance of a batch per blockIdx.z
template <class S>
__global__ void __launch_bounds__(256, 2) k_zsolve64p_b(const ZJob* __restrict__ jobs) {
  const ZJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_zsolve64p_body<S>(j.g, j.a);
}

// ---------------------------------------------------------------------------
// nz = 24, 32, 48, 4 <= Lmax <= 7: the half-tile split kernel with a wave-local transform
// ---------------------------------------------------------------------------
// Same workgroup shape as k_zsolve64p (four waves: two columns x two degree parts, two
// workgroups per CU; the halves of a tile are blocks b and b + 8 on one XCD), lane = z for
// lanes 0..NZ-1 (the others hold zeros through the solve).  The NZ-point transform of a
// batch of eight harmonics is a four-step N1 x 8 (N1 = NZ / 8 = 3, 4, 6) in the wave's LDS
// region: with n = j + N1 m and k = k1 + 8 k2, a DFT-8 over m for each (harmonic, j) (NZ
// lanes), the twiddle W_NZ^(j k1), then a DFT-N1 over j for each (harmonic, k1) (64 lanes).
// The generic kernel it replaces reads 16 of every 64 bytes of a tile per thread and runs
// runtime-planned block-wide Stockham passes.
template <int NZ>
struct HalfTile {
  static constexpr int N1 = NZ / 8;
  static constexpr int kPad = NZ + NZ / 8;  // padded units per harmonic (pad8(NZ - 1) < kPad)
  static constexpr int kRegion = (kHB * kPad + 2) > (kXch + 2) ? (kHB * kPad + 2) : (kXch + 2);
};

// X[k] = sum_n x[n] exp(DIR 2 pi i n k / 6): two DFT-3 and one radix-2 step
template <int DIR>
__device__ __forceinline__ void dft6(double2 (&v)[6]) {
  double2 e[3] = {v[0], v[2], v[4]};
  double2 o[3] = {v[1], v[3], v[5]};
  Dft<3, DIR>::run(e);
  Dft<3, DIR>::run(o);
  const double s = 0.86602540378443864676;  // sqrt(3)/2
  // W6^k = exp(DIR 2 pi i k / 6): W6 = (1/2, DIR s), W6^2 = (-1/2, DIR s)
  const double2 o1 = make_double2(0.5 * o[1].x - DIR * s * o[1].y, 0.5 * o[1].y + DIR * s * o[1].x);
  const double2 o2 = make_double2(-0.5 * o[2].x - DIR * s * o[2].y, -0.5 * o[2].y + DIR * s * o[2].x);
  v[0] = cadd(e[0], o[0]);
  v[3] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[4] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[5] = csub(e[2], o2);
}

template <int N, int DIR>
__device__ __forceinline__ void small_dft(double2 (&v)[N]) {
  if constexpr (N == 6) dft6<DIR>(v);
  else Dft<N, DIR>::run(v);
}

// the NZ-point transforms of the eight harmonics in region wl (element z of slot h at
// wl[h kPad + pad8(z)]), in place; DIR < 0 forward, DIR > 0 inverse (unnormalised)
template <int NZ, int DIR>
__device__ __forceinline__ void zh_batch(double2* wl, int lane, const double2* __restrict__ twz) {
  constexpr int N1 = HalfTile<NZ>::N1, P = HalfTile<NZ>::kPad;
  if (lane < NZ) {
    const int h = lane / N1, j = lane - N1 * (lane / N1);
    double2 u[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) u[m] = wl[h * P + pad8(j + N1 * m)];
    Dft<8, DIR>::run(u);
#pragma unroll
    for (int k1 = 1; k1 < 8; ++k1) u[k1] = twiddle<DIR>(u[k1], twz[j * k1]);
    // the positions j + N1 k1 this lane writes are the ones it read
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) wl[h * P + pad8(j + N1 * k1)] = u[k1];
  }
  wave_sync();
  {
    const int h = lane >> 3, k1 = lane & 7;
    double2 u[N1];
#pragma unroll
    for (int j = 0; j < N1; ++j) u[j] = wl[h * P + pad8(j + N1 * k1)];
    small_dft<N1, DIR>(u);
    wave_sync();  // every lane has read its inputs before any output overwrites them
#pragma unroll
    for (int k2 = 0; k2 < N1; ++k2) wl[h * P + pad8(k1 + 8 * k2)] = u[k2];
  }
  wave_sync();
}

template <int NZ, class S>
__device__ __forceinline__ void k_zsolvehp_body(const Geom& g, const SolveArgs& a) {
  constexpr int kPadZ = HalfTile<NZ>::kPad, kRegionZ = HalfTile<NZ>::kRegion;
  extern __shared__ double2 lds[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int part = wv >> 1, c = wv & 1, gi = threadIdx.x & 127;
  const int ny = g.ny, nxh = g.nxh, nt = g.nt;
  const int nb = gridDim.x, b = blockIdx.x;
  int tl, half;
  if ((nb & 15) == 0) {
    const int t = b >> 3;
    half = t & 1;
    tl = (t >> 1) * 8 + (b & 7);
  } else {
    half = b & 1;
    tl = b >> 1;
  }
  const int t4 = tl % nt, ky = tl / nt;
  if (t4 * kKxTile + 2 * half >= nxh) return;  // both columns are padding (uniform)
  const int kx = t4 * kKxTile + 2 * half + c;
  const bool real_col = kx < nxh;  // wave-uniform
  // thread gi of a part moves element (z = gi / 2, column gi % 2) of each harmonic
  const int tz = gi >> 1, tc = gi & 1;
  const bool io = tz < NZ && t4 * kKxTile + 2 * half + tc < nxh;
  const int go = kKxTile * tz + tc;
  const bool active = a.solve && real_col;
  const long long istride = spec_index(1, 0, 0, 0, ny, nt, NZ);
  const int nh = part ? PartHi::NH : PartLo::NH;
  const int hb0 = part ? PartHi::B0 : PartLo::B0;
  S* hp = reinterpret_cast<S*>(a.spec) + spec_index(0, ky, t4 * kKxTile, 0, ny, nt, NZ) + 2 * half +
          (long long)hb0 * istride;
  double2* glds = lds + part * (2 * kRegionZ);
  double2* wl = glds + c * kRegionZ;
  double2* xch = lds + c * kRegionZ;  // the Lo wave's region of column c
  const double2* twz = g.az.tw;        // exp(-2 pi i t / NZ)
  const int zl = lane < NZ ? lane : 0;
  const KFrame f(signed_k(kx, g.nx) * a.kfac[0], signed_k(ky, ny) * a.kfac[1],
                 signed_k(zl, NZ) * a.kfac[2], a.sigma, a.ds);
  double2 v[kNV];
#pragma unroll
  for (int q = 0; q < kNV; ++q) v[q] = make_double2(0.0, 0.0);
  // forward: per batch of eight harmonics, the part's tiles into LDS (the next batch's loads
  // in flight), the 48-point transforms, lane = kz back into registers
  double2 nxt[kHB];
#pragma unroll
  for (int q = 0; q < kHB; ++q)
    nxt[q] = (io && q < nh) ? spec_ld(hp + (long long)q * istride + go) : make_double2(0.0, 0.0);
#pragma unroll
  for (int bb = 0; bb < kNBt; ++bb) {
    const int b0 = bb * kHB;
    __syncthreads();  // every wave of the part has read the previous batch out of its region
    if (tz < NZ) {
#pragma unroll
      for (int q = 0; q < kHB; ++q)
        if (b0 + q < nh) glds[tc * kRegionZ + q * kPadZ + pad8(tz)] = nxt[q];
    }
    if (bb + 1 < kNBt) {
#pragma unroll
      for (int q = 0; q < kHB; ++q) {
        const int hq = b0 + kHB + q;
        nxt[q] = (io && hq < nh) ? spec_ld(hp + (long long)hq * istride + go) : make_double2(0.0, 0.0);
      }
    }
    __syncthreads();
    if (real_col && b0 < nh) {
      zh_batch<NZ, -1>(wl, lane, twz);
      if (lane < NZ) {
#pragma unroll
        for (int q = 0; q < kHB; ++q)
          if (b0 + q < kNV && b0 + q < nh) v[b0 + q] = wl[q * kPadZ + pad8(lane)];
      }
    }
  }
  if (active) {
    if (part == 0) {
      PartLo::to_frame(v, f);
      tri_lower_forward<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
    } else {
      PartHi::to_frame(v, f);
    }
  }
  __syncthreads();
  if (part == 1 && active) tri_upper<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
  __syncthreads();
  if (active) {
    if (part == 0) {
      tri_lower_back<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
      PartLo::from_frame(v, f);
    } else {
      PartHi::from_frame(v, f);
    }
  }
  // inverse: per batch, lane = kz into LDS, the inverse transforms, the tiles out
#pragma unroll
  for (int bb = 0; bb < kNBt; ++bb) {
    const int b0 = bb * kHB;
    __syncthreads();  // the previous batch's tiles (and the crossing chains) have been read
    if (real_col && b0 < nh) {
      if (lane < NZ) {
#pragma unroll
        for (int q = 0; q < kHB; ++q)
          if (b0 + q < kNV && b0 + q < nh) wl[q * kPadZ + pad8(lane)] = v[b0 + q];
      }
      wave_sync();
      zh_batch<NZ, +1>(wl, lane, twz);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kHB; ++q)
      if (io && b0 + q < nh)
        spec_st(hp + (long long)(b0 + q) * istride + go, glds[tc * kRegionZ + q * kPadZ + pad8(tz)]);
  }
}

template <int NZ, class S>
__global__ void __launch_bounds__(256, 2) k_zsolvehp(Geom g, SolveArgs a) {
  k_zsolvehp_body<NZ, S>(g, a);
}

// the same for one instance of a batch per blockIdx.z
template <int NZ, class S>
__global__ void __launch_bounds__(256, 2) k_zsolvehp_b(const ZJob* __restrict__ jobs) {
  const ZJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_zsolvehp_body<NZ, S>(j.g, j.a);
}

// ---------------------------------------------------------------------------
// nz = 128: two waves per (kx, ky) column
// ---------------------------------------------------------------------------
// Wave p of a column works on z = m + 64 p (lane m).  Forward: one radix-2 decimation-
// in-frequency step, a = x(m) + x(m + 64) in wave 0, b = (x(m) - x(m + 64)) w^m in wave 1
// with w = exp(-2 pi i / 128), then the wave-local 64-point transform, after which lane m
// of wave p holds kz = 2 m + p for all harmonics (the solve needs nothing else).  Inverse:
// wave-local 64-point inverse, then x(m) = a'(m) + w^-m b'(m), x(m + 64) = a'(m) - w^-m b'(m).
// k_zsolve128 (workgroup = two adjacent columns of a tile) with the FP64 spectrum moves the
// columns between HBM and LDS with coalesced accesses and forms the radix-2 step from LDS
// (z128_load_dif, z128_dit_store); with the FP32 spectrum, and in k_zsolve128s (one column
// per workgroup), each wave loads its half directly and the halves are exchanged through
// LDS (z128_exchange).
constexpr int kXB = 16;  // harmonics per cross-wave exchange
constexpr int kRegion128 = kRegion > kXB * 64 ? kRegion : kXB * 64;

// exchange with the partner wave of the column: own values in, partner's values out,
// kXB harmonics at a time; every wave of the workgroup takes part in the barriers
// NLOOP >= NH sets the number of batches, so that waves exchanging different numbers of
// harmonics in one workgroup still pass the same number of barriers
template <int DIR, int NH, int NLOOP = NH, int XB = kXB, int N>
__device__ __forceinline__ void z128_exchange(double2 (&v)[N], double2* own, const double2* other,
                                              int lane, int p, double2 t, bool real_col) {
  static_assert(XB <= kXB, "exchange batch fits a wave region");
  static_assert(NLOOP >= NH, "every harmonic is exchanged");
#pragma unroll
  for (int b0 = 0; b0 < NLOOP; b0 += XB) {
    __syncthreads();  // the partner has read the previous batch
    if (real_col) {
#pragma unroll
      for (int q = 0; q < XB; ++q)
        if (b0 + q < NH) own[q * 64 + lane] = v[b0 + q];
    }
    __syncthreads();
    if (real_col) {
#pragma unroll
      for (int q = 0; q < XB; ++q) {
        if (b0 + q < NH) {
          const double2 o = other[q * 64 + lane], x = v[b0 + q];
          if (DIR < 0) {
            // forward: wave 0 keeps x(m) + x(m + 64), wave 1 (x(m) - x(m + 64)) w^m
            v[b0 + q] = p == 0 ? make_double2(x.x + o.x, x.y + o.y)
                               : twiddle<-1>(make_double2(o.x - x.x, o.y - x.y), t);
          } else {
            // inverse: wave 1 arrives holding w^-m b'(m)
            v[b0 + q] = p == 0 ? make_double2(x.x + o.x, x.y + o.y)
                               : make_double2(o.x - x.x, o.y - x.y);
          }
        }
      }
    }
  }
}

// The two columns of a k_zsolve128 workgroup are adjacent in the tile: element (h, z, cc)
// of the pair at h * istride + 4 z + cc.  Thread t moves element (z = t >> 1, cc = t & 1)
// of each harmonic (32 contiguous bytes per z for the pair), kPB harmonics per batch
// through LDS rows of kPC entries (132: the two columns' rows fall on distinct banks).
constexpr int kPB = 15;
constexpr int kPC = 132;
static_assert(kPB * 2 * kPC <= 4 * kRegion128, "pair batch fits the workgroup's LDS");

// forward: every harmonic of the pair loaded at once (thread-owned elements land in v),
// then per batch through LDS into the radix-2 decimation-in-frequency step: wave p of
// column c gets x(m) + x(m + 64) (p = 0) or (x(m) - x(m + 64)) w^m (p = 1)
template <class S>
__device__ __forceinline__ void z128_load_dif(double2 (&v)[kM], double2* buf, const S* __restrict__ pair,
                                              long long istride, int c, int p, int lane, double2 t,
                                              bool real_col, bool io) {
  const int tz = threadIdx.x >> 1, tc = threadIdx.x & 1;
#pragma unroll
  for (int h = 0; h < kM; ++h)
    v[h] = io ? spec_ld(pair + (long long)h * istride + 4 * tz + tc) : make_double2(0.0, 0.0);
#pragma unroll
  for (int b0 = 0; b0 < kM; b0 += kPB) {
    __syncthreads();  // the previous batch has been read
#pragma unroll
    for (int q = 0; q < kPB; ++q)
      if (b0 + q < kM) buf[(q * 2 + tc) * kPC + tz] = v[b0 + q];
    __syncthreads();
    if (real_col) {
#pragma unroll
      for (int q = 0; q < kPB; ++q) {
        if (b0 + q < kM) {
          const double2 x0 = buf[(q * 2 + c) * kPC + lane], x1 = buf[(q * 2 + c) * kPC + lane + 64];
          v[b0 + q] = p == 0 ? make_double2(x0.x + x1.x, x0.y + x1.y)
                             : twiddle<-1>(make_double2(x0.x - x1.x, x0.y - x1.y), t);
        }
      }
    }
  }
}

// inverse: wave p of column c holds a'(m) (p = 0) or w^-m b'(m) (p = 1); per batch through
// LDS, thread (z, cc) forms x(z) = a' + w^-m b' (z = m < 64) or a' - w^-m b' (z = m + 64)
// and stores it
template <class S>
__device__ __forceinline__ void z128_dit_store(const double2 (&v)[kM], double2* buf, S* __restrict__ pair,
                                               long long istride, int c, int p, int lane,
                                               bool real_col, bool io) {
  const int tz = threadIdx.x >> 1, tc = threadIdx.x & 1, m = tz & 63;
#pragma unroll
  for (int b0 = 0; b0 < kM; b0 += kPB) {
    __syncthreads();  // the previous batch has been read
    if (real_col) {
#pragma unroll
      for (int q = 0; q < kPB; ++q)
        if (b0 + q < kM) buf[(q * 2 + c) * kPC + 64 * p + lane] = v[b0 + q];
    }
    __syncthreads();
    if (io) {
#pragma unroll
      for (int q = 0; q < kPB; ++q) {
        if (b0 + q < kM) {
          const double2 a0 = buf[(q * 2 + tc) * kPC + m], b1 = buf[(q * 2 + tc) * kPC + 64 + m];
          spec_st(pair + (long long)(b0 + q) * istride + 4 * tz + tc,
                  tz < 64 ? make_double2(a0.x + b1.x, a0.y + b1.y)
                          : make_double2(a0.x - b1.x, a0.y - b1.y));
        }
      }
    }
  }
}

template <int OCC, class S>
__device__ __forceinline__ void k_zsolve128_body(const Geom& g, const SolveArgs& a) {
  __shared__ double2 lds[4 * kRegion128];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int c = wv >> 1, p = wv & 1;
  const int ny = g.ny, nxh = g.nxh, nt = g.nt;
  // block -> (tile, half): blocks b and b + 8 (consecutive on XCD b mod 8) share a tile
  const int nb = gridDim.x, b = blockIdx.x;
  int tl, half;
  if ((nb & 15) == 0) {
    const int t = b >> 3;
    half = t & 1;
    tl = (t >> 1) * 8 + (b & 7);
  } else {
    half = b & 1;
    tl = b >> 1;
  }
  const int t4 = tl % nt, ky = tl / nt;
  if (t4 * kKxTile + 2 * half >= nxh) return;  // both columns are padding (uniform)
  const int kxl = 2 * half + c;
  const int kx = t4 * kKxTile + kxl;
  const bool real_col = kx < nxh;                                     // wave-uniform
  const bool io = t4 * kKxTile + 2 * half + (threadIdx.x & 1) < nxh;  // the column this thread moves
  const long long istride = spec_index(1, 0, 0, 0, ny, nt, 128);
  // the workgroup's column pair kx = 4 t4 + 2 half + {0, 1}: element (i, z, cc) at
  // i * istride + 4 z + cc
  S* pair = reinterpret_cast<S*>(a.spec) + spec_index(0, ky, t4 * kKxTile, 0, ny, nt, 128) + 2 * half;
  double2* wl = lds + wv * kRegion128;
  double2 tw[8];
  {
    const int j = lane & 7;
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) tw[k2] = g.az.tw[2 * j * k2];  // exp(-2 pi i j k2 / 64)
  }
  const double2 t = g.az.tw[lane];  // exp(-2 pi i m / 128)
  double2 v[kM];
  // FP64 spectrum: the column pair through LDS with coalesced accesses; FP32 spectrum
  // (8-byte elements): direct loads of each wave's half and an exchange between the waves,
  // which is faster there (128^3/L8: 1.9 vs 2.4 ms)
  constexpr bool kPair = std::is_same<S, double2>::value;
  [[maybe_unused]] S* col = pair + c + 64 * 4 * p;  // this wave's half column: (i, m) at i * istride + 4 m
  if constexpr (kPair) {
    z128_load_dif(v, lds, pair, istride, c, p, lane, t, real_col, io);
  } else {
    if (real_col) {
#pragma unroll
      for (int h = 0; h < kM; ++h) v[h] = spec_ld(col + (long long)h * istride + 4 * lane);
    }
    z128_exchange<-1, kM>(v, wl, lds + (wv ^ 1) * kRegion128, lane, p, t, real_col);
  }
  __syncthreads();  // the LDS rows are reused by the 64-point transforms
  if (real_col) z64_fwd_regs<kM, (kL >= 7)>(v, wl, lane, tw);
  if (a.solve && real_col) {
    const double kxv = signed_k(kx, g.nx) * a.kfac[0];
    const double kyv = signed_k(ky, ny) * a.kfac[1];
    const double kzv = signed_k(2 * lane + p, 128) * a.kfac[2];
    Solver<kL>::solve(v, kxv, kyv, kzv, a.cdelta, a.sigma, a.ds, a.dslam);
  }
  if (real_col) {
    z64_inv_regs<kM, (kL >= 7)>(v, wl, lane, tw);
    if (p == 1) {
#pragma unroll
      for (int h = 0; h < kM; ++h) v[h] = twiddle<+1>(v[h], t);
    }
  }
  if constexpr (kPair) {
    z128_dit_store(v, lds, pair, istride, c, p, lane, real_col, io);
  } else {
    z128_exchange<+1, kM>(v, wl, lds + (wv ^ 1) * kRegion128, lane, p, t, real_col);
    if (real_col) {
#pragma unroll
      for (int h = 0; h < kM; ++h) spec_st(col + (long long)h * istride + 4 * lane, v[h]);
    }
  }
}

template <int OCC, class S>
__global__ void __launch_bounds__(256, OCC) k_zsolve128(Geom g, SolveArgs a) {
  k_zsolve128_body<OCC, S>(g, a);
}

// the same for one instance of a batch per blockIdx.z
template <int OCC, class S>
__global__ void __launch_bounds__(256, OCC) k_zsolve128_b(const ZJob* __restrict__ jobs) {
  const ZJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_zsolve128_body<OCC, S>(j.g, j.a);
}

// nz = 128 with the degree split of k_zsolve64s: workgroup = one column, four waves
// wv = 2 part + p (part 0: l <= kLS, part 1: l > kLS; p: z half).  The radix-2 step pairs
// wv with wv ^ 1, the crossing tridiagonal chains pair wv with wv ^ 2 through the region of
// the part-0 wave of the same z half.  Each wave holds about half the harmonics, so two
// waves fit per SIMD at Lmax 7 and 8.
template <int OCC, class S>
__device__ __forceinline__ void k_zsolve128s_body(const Geom& g, const SolveArgs& a) {
  __shared__ double2 lds[4 * kRegion128];
  static_assert(kXch + 2 <= kRegion128, "crossing chains fit a wave region");
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int p = wv & 1, part = wv >> 1;
  const int ny = g.ny, nxh = g.nxh, nt = g.nt;
  // block -> column: the four columns of a 4-kx tile run on one XCD (blocks b, b + 8,
  // b + 16, b + 24), so each tile's cache lines are fetched once
  const int nb = gridDim.x, b = blockIdx.x;
  int tl, kxl;
  if ((nb & 31) == 0) {
    const int t = b >> 3;
    kxl = t & 3;
    tl = (t >> 2) * 8 + (b & 7);
  } else {
    kxl = b & 3;
    tl = b >> 2;
  }
  const int t4 = tl % nt, ky = tl / nt;
  const int kx = t4 * kKxTile + kxl;
  if (kx >= nxh) return;  // a padding column (uniform over the workgroup)
  const long long istride = spec_index(1, 0, 0, 0, ny, nt, 128);
  S* col = reinterpret_cast<S*>(a.spec) + spec_index(0, ky, t4 * kKxTile, 64 * p, ny, nt, 128) + kxl;
  double2* wl = lds + wv * kRegion128;
  const double2* partner = lds + (wv ^ 1) * kRegion128;
  double2* xch = lds + p * kRegion128;  // region of the part-0 wave of this z half
  // the twiddles are reloaded (cache hits) around each transform instead of being held
  // through the solve, where the registers are needed for the harmonics
  const double2* twz = g.az.tw;  // exp(-2 pi i m / 128); exp(-2 pi i j k2 / 64) = twz[2 j k2]
  const bool active = a.solve;
  const KFrame f(signed_k(kx, g.nx) * a.kfac[0], signed_k(ky, ny) * a.kfac[1],
                 signed_k(2 * lane + p, 128) * a.kfac[2], a.sigma, a.ds);
  double2 v[kNV];
  {
    const double2 t = g.az.tw[lane];  // exp(-2 pi i m / 128)
    if (part == 0) {
#pragma unroll
      for (int h = 0; h < PartLo::NH; ++h) v[h] = spec_ld(col + (long long)(PartLo::B0 + h) * istride + 4 * lane);
      z128_exchange<-1, PartLo::NH, kNV>(v, wl, partner, lane, p, t, true);
    } else {
#pragma unroll
      for (int h = 0; h < PartHi::NH; ++h) v[h] = spec_ld(col + (long long)(PartHi::B0 + h) * istride + 4 * lane);
      z128_exchange<-1, PartHi::NH, kNV>(v, wl, partner, lane, p, t, true);
    }
  }
  __syncthreads();  // the exchange regions are reused by the transforms and the chains
  {
    double2 tw[8];
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) tw[k2] = twz[2 * (lane & 7) * k2];
    if (part == 0) {
      z64_fwd_regs<PartLo::NH>(v, wl, lane, tw);
    } else {
      z64_fwd_regs<PartHi::NH>(v, wl, lane, tw);
    }
  }
  // xch is the part-0 wave's own region: its transform is done (program order), and the
  // part-1 wave reads xch only after the next barrier
  if (part == 0) {
    if (active) {
      PartLo::to_frame(v, f);
      tri_lower_forward<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
    }
  } else {
    if (active) PartHi::to_frame(v, f);
  }
  __syncthreads();
  if (part == 1 && active) tri_upper<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
  __syncthreads();
  if (part == 0) {
    if (active) {
      tri_lower_back<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
      PartLo::from_frame(v, f);
    }
  } else {
    if (active) PartHi::from_frame(v, f);
  }
  // after the last barrier only the part-0 wave reads xch (its own region), so the
  // transforms below need no further barrier
  double2 tw[8];
#pragma unroll
  for (int k2 = 0; k2 < 8; ++k2) tw[k2] = twz[2 * (lane & 7) * k2];
  const double2 t = g.az.tw[lane];
  if (part == 0) {
    z64_inv_regs<PartLo::NH>(v, wl, lane, tw);
    if (p == 1) {
#pragma unroll
      for (int h = 0; h < PartLo::NH; ++h) v[h] = twiddle<+1>(v[h], t);
    }
    z128_exchange<+1, PartLo::NH, kNV>(v, wl, partner, lane, p, t, true);
#pragma unroll
    for (int h = 0; h < PartLo::NH; ++h) spec_st(col + (long long)(PartLo::B0 + h) * istride + 4 * lane, v[h]);
  } else {
    z64_inv_regs<PartHi::NH>(v, wl, lane, tw);
    if (p == 1) {
#pragma unroll
      for (int h = 0; h < PartHi::NH; ++h) v[h] = twiddle<+1>(v[h], t);
    }
    z128_exchange<+1, PartHi::NH, kNV>(v, wl, partner, lane, p, t, true);
#pragma unroll
    for (int h = 0; h < PartHi::NH; ++h) spec_st(col + (long long)(PartHi::B0 + h) * istride + 4 * lane, v[h]);
  }
}

template <int OCC, class S>
__global__ void __launch_bounds__(256, OCC) k_zsolve128s(Geom g, SolveArgs a) {
  k_zsolve128s_body<OCC, S>(g, a);
}

// the same for one instance of a batch per blockIdx.z
template <int OCC, class S>
__global__ void __launch_bounds__(256, OCC) k_zsolve128s_b(const ZJob* __restrict__ jobs) {
  const ZJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_zsolve128s_body<OCC, S>(j.g, j.a);
}

// generic kernel B (nz other than 64 and 128, or any nz with the generic kernel forced),
// 4 <= Lmax <= 7, with the harmonics of each
// k-point split by degree over two threads: thread t < TX nz holds l <= kLS of k-point
// (tx, z), thread t + TX nz the same k-point's l > kLS, which doubles the threads of a
// small grid and halves their registers.  The z transforms of a batch of eight harmonics
// run over the 2 TX column-parts as one block-wide Stockham transform (element z of
// column-part j of batch slot h at LDS (h 2 TX + j) nz + z, i.e. thread t's element at
// h blockDim + t); the crossing chains are exchanged through LDS as in k_zsolve64s, one
// region per 64 k-points.
template <int N, int DIR>
__device__ __forceinline__ void z_transform_n(double2 (&v)[N], double2* lds, const AxisDev& az,
                                              int ncols, int lidx) {
  const int nz = az.n;
#pragma unroll
  for (int b0 = 0; b0 < N; b0 += kHB) {
    const int nb = (N - b0) < kHB ? (N - b0) : kHB;
#pragma unroll
    for (int h = 0; h < kHB; ++h)
      if (b0 + h < N) lds[h * ncols * nz + lidx] = v[b0 + h];
    __syncthreads();
    fft_axis<DIR, kHB>(lds, az, nb * ncols, 1, nz);
#pragma unroll
    for (int h = 0; h < kHB; ++h)
      if (b0 + h < N) v[b0 + h] = lds[h * ncols * nz + lidx];
    __syncthreads();
  }
}

template <class S>
__device__ __forceinline__ void k_zsolve_split_body(const Geom& g, const SolveArgs& a) {
  extern __shared__ double2 lds[];
  const int nz = g.nz, ny = g.ny, nxh = g.nxh;
  const int TX = a.TX;
  const int np = TX * nz;  // threads per part; 2 TX nz == blockDim.x
  const int t = threadIdx.x;
  const int part = t >= np ? 1 : 0;
  const int li = t - part * np;
  const int tx = li / nz, z = li - tx * nz;
  const int kx = blockIdx.x * TX + tx, ky = blockIdx.y;
  const bool valid = kx < nxh;
  const bool active = a.solve && valid;
  const int nh = part ? PartHi::NH : PartLo::NH;
  const int hb0 = part ? PartHi::B0 : PartLo::B0;
  const long long istride = spec_index(1, 0, 0, 0, ny, g.nt, nz);
  S* p = reinterpret_cast<S*>(a.spec) + spec_index(0, ky, valid ? kx : 0, z, ny, g.nt, nz) +
         (long long)hb0 * istride;
  double2 v[kNV];
#pragma unroll
  for (int q = 0; q < kNV; ++q)
    v[q] = (valid && q < nh) ? spec_ld(p + q * istride) : make_double2(0.0, 0.0);
  z_transform_n<kNV, -1>(v, lds, g.az, 2 * TX, t);
  // the transform's rows are free after its last barrier: crossing-chain regions
  double2* xch = lds + (li >> 6) * kXch;
  const int lane = li & 63;
  const KFrame f(signed_k(kx, g.nx) * a.kfac[0], signed_k(ky, ny) * a.kfac[1],
                 signed_k(z, nz) * a.kfac[2], a.sigma, a.ds);
  if (active) {
    if (part == 0) {
      PartLo::to_frame(v, f);
      tri_lower_forward<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
    } else {
      PartHi::to_frame(v, f);
    }
  }
  __syncthreads();
  if (part == 1 && active) tri_upper<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
  __syncthreads();
  if (active) {
    if (part == 0) {
      tri_lower_back<kL, kLS>(v, a.cdelta, a.dslam, f.gk, xch, lane);
      PartLo::from_frame(v, f);
    } else {
      PartHi::from_frame(v, f);
    }
  }
  __syncthreads();  // every crossing chain has been read before the transform reuses LDS
  z_transform_n<kNV, +1>(v, lds, g.az, 2 * TX, t);
  if (valid) {
#pragma unroll
    for (int q = 0; q < kNV; ++q)
      if (q < nh) spec_st(p + q * istride, v[q]);
  }
}

template <class S>
__global__ void __launch_bounds__(512) k_zsolve_split(Geom g, SolveArgs a) {
  k_zsolve_split_body<S>(g, a);
}

// the same for one instance of a batch per blockIdx.z
template <class S>
__global__ void __launch_bounds__(512) k_zsolve_split_b(const ZJob* __restrict__ jobs) {
  const ZJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_zsolve_split_body<S>(j.g, j.a);
}

}  // namespace per_l_<L>

using namespace WLC_CAT(per_l_, WLC_L);

namespace {
// the kernel-B variant for (g, a): single-launch kernel, its batched twin, grid, block, LDS
struct ZPlan {
  const void* k1 = nullptr;
  const void* kb = nullptr;
  dim3 grid, block;
  size_t lds = 0;
};

template <class S>
hipError_t zsolve_plan(const Geom& g, const SolveArgs& a, ZPlan* z) {
  if (a.M != kM) return hipErrorInvalidValue;
  if constexpr (std::is_same<S, double2>::value) {
    // one workgroup of eight waves per 4-kx tile; FP64 spectrum only: with the FP32 half
    // spectrum (2 KB tiles) the two-column kernel below is faster (64^3/L6: 69 vs 93 us)
    if ((g.nz == 24 || g.nz == 32 || g.nz == 48) && !a.generic && kL >= 4 && kL <= 7 && a.split == 3) {
      // two workgroups of four waves per 4-kx tile, wave-local NZ-point transforms
      z->grid = dim3(2 * g.nt * g.ny);
      z->block = dim3(256);
      if (g.nz == 24) {
        z->k1 = (const void*)k_zsolvehp<24, S>;
        z->kb = (const void*)k_zsolvehp_b<24, S>;
        z->lds = sizeof(double2) * (size_t)(2 * 2 * HalfTile<24>::kRegion);
      } else if (g.nz == 32) {
        z->k1 = (const void*)k_zsolvehp<32, S>;
        z->kb = (const void*)k_zsolvehp_b<32, S>;
        z->lds = sizeof(double2) * (size_t)(2 * 2 * HalfTile<32>::kRegion);
      } else {
        z->k1 = (const void*)k_zsolvehp<48, S>;
        z->kb = (const void*)k_zsolvehp_b<48, S>;
        z->lds = sizeof(double2) * (size_t)(2 * 2 * HalfTile<48>::kRegion);
      }
      return hipSuccess;
    }
    if (g.nz == 64 && !a.generic && kL >= 5 && kL <= 7 && a.split == 3) {
      // two workgroups of four waves per 4-kx tile (half a tile each)
      z->k1 = (const void*)k_zsolve64p<S>;
      z->kb = (const void*)k_zsolve64p_b<S>;
      z->grid = dim3(2 * g.nt * g.ny);
      z->block = dim3(256);
      z->lds = sizeof(double2) * (size_t)(2 * 2 * kRegionS);
      return hipSuccess;
    }
    if (g.nz == 64 && !a.generic && kL >= 5 && kL <= 7 && a.split == 2) {
      z->k1 = (const void*)k_zsolve64t<S>;
      z->kb = (const void*)k_zsolve64t_b<S>;
      z->grid = dim3(g.nt * g.ny);
      z->block = dim3(512);
      z->lds = sizeof(double2) * (size_t)kTileLds;
      return hipSuccess;
    }
  }
  if (g.nz == 64 && !a.generic && kL >= 5 && kL <= 7 && a.split) {
    // two workgroups per 4-kx tile
    z->k1 = (const void*)k_zsolve64s<2, S>;
    z->kb = (const void*)k_zsolve64s_b<2, S>;
    z->grid = dim3(2 * g.nt * g.ny);
    z->block = dim3(256);
    return hipSuccess;
  }
  if (g.nz == 128 && !a.generic && kL >= 5 && a.split) {
    // one workgroup (one column, harmonics split over two waves per z half) per column; two
    // waves per SIMD up to Lmax 7, one at Lmax 8 (the 45 harmonics of l = 6..8 need more
    // than half the register file)
    z->k1 = (const void*)k_zsolve128s<(kL >= 8 ? 1 : 2), S>;
    z->kb = (const void*)k_zsolve128s_b<(kL >= 8 ? 1 : 2), S>;
    z->grid = dim3(4 * g.nt * g.ny);
    z->block = dim3(256);
    return hipSuccess;
  }
  if (g.nz == 128 && !a.generic) {
    // two workgroups (two columns each) per 4-kx tile
    z->k1 = (const void*)k_zsolve128<(kL >= 7 ? 1 : 2), S>;
    z->kb = (const void*)k_zsolve128_b<(kL >= 7 ? 1 : 2), S>;
    z->grid = dim3(2 * g.nt * g.ny);
    z->block = dim3(256);
    return hipSuccess;
  }
  if (g.nz == 64 && !a.generic) {
    static_assert(kKxTile == 4, "one wave per kx of a tile");
    z->k1 = a.occ == 2 ? (const void*)k_zsolve64<2, S> : (const void*)k_zsolve64<1, S>;
    z->kb = a.occ == 2 ? (const void*)k_zsolve64_b<2, S> : (const void*)k_zsolve64_b<1, S>;
    z->grid = dim3(g.nt, g.ny);
    z->block = dim3(256);
    return hipSuccess;
  }
  if (a.TX * g.nz > 256 || a.TX < 1) return hipErrorInvalidValue;
  if constexpr (kL >= 4 && kL <= 7) {
    if (a.split) {
      // two threads per k-point (degree split): 2 TX nz threads
      const int np = a.TX * g.nz;
      const size_t rows = (size_t)kHB * 2 * np, chains = (size_t)((np + 63) / 64) * kXch;
      z->k1 = (const void*)k_zsolve_split<S>;
      z->kb = (const void*)k_zsolve_split_b<S>;
      z->grid = dim3((g.nxh + a.TX - 1) / a.TX, g.ny);
      z->block = dim3(2 * np);
      z->lds = sizeof(double2) * (rows > chains ? rows : chains);
      return hipSuccess;
    }
  }
  z->k1 = (const void*)k_zsolve<S>;
  z->kb = (const void*)k_zsolve_b<S>;
  z->grid = dim3((g.nxh + a.TX - 1) / a.TX, g.ny);
  z->block = dim3(a.TX * g.nz);
  z->lds = sizeof(double2) * (size_t)kHB * a.TX * g.nz;
  return hipSuccess;
}

hipError_t plan_of(const Geom& g, const SolveArgs& a, ZPlan* z) {
  return g.spec32 ? zsolve_plan<float2>(g, a, z) : zsolve_plan<double2>(g, a, z);
}
}  // namespace

hipError_t WLC_CAT(zsolve_launch_L, WLC_L)(const Geom& g, const SolveArgs& a, hipStream_t st) {
  ZPlan z;
  hipError_t e = plan_of(g, a, &z);
  if (e != hipSuccess) return e;
  if (z.lds) {
    e = hipFuncSetAttribute(z.k1, hipFuncAttributeMaxDynamicSharedMemorySize, (int)z.lds);
    if (e != hipSuccess) return e;
  }
  Geom gg = g;
  SolveArgs aa = a;
  void* args[] = {&gg, &aa};
  return hipLaunchKernel(z.k1, z.grid, z.block, args, z.lds, st);
}

hipError_t WLC_CAT(zsolve_batch_L, WLC_L)(const ZJob& rep, const ZJob* jobs, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  ZPlan z;
  hipError_t e = plan_of(rep.g, rep.a, &z);
  if (e != hipSuccess) return e;
  if (z.lds) {
    e = hipFuncSetAttribute(z.kb, hipFuncAttributeMaxDynamicSharedMemorySize, (int)z.lds);
    if (e != hipSuccess) return e;
  }
  const ZJob* jp = jobs;
  void* args[] = {&jp};
  return hipLaunchKernel(z.kb, dim3(z.grid.x, z.grid.y, n), z.block, args, z.lds, st);
}

LaunchKey WLC_CAT(zsolve_key_L, WLC_L)(const Geom& g, const SolveArgs& a) {
  ZPlan z;
  if (plan_of(g, a, &z) != hipSuccess) return LaunchKey{nullptr, 0, 0, 0, 0};
  return LaunchKey{z.k1, z.grid.x, z.grid.y, z.block.x, z.lds};
}

namespace WLC_CAT(per_l_, WLC_L) {
__global__ void k_test_solve(int nk, const double* kvec, double cdelta, double sigma, double ds,
                             double dslam, double2* v) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nk) return;
  double2 x[kM];
#pragma unroll
  for (int i = 0; i < kM; ++i) x[i] = v[(long long)p * kM + i];
  Solver<kL>::solve(x, kvec[3 * p], kvec[3 * p + 1], kvec[3 * p + 2], cdelta, sigma, ds, dslam);
#pragma unroll
  for (int i = 0; i < kM; ++i) v[(long long)p * kM + i] = x[i];
}

}  // namespace per_l_<L>

hipError_t WLC_CAT(test_solve_L, WLC_L)(int nk, const double* kvec, double cdelta, double sigma,
                                        double ds, double dslam, double2* v, hipStream_t st) {
  hipLaunchKernelGGL(k_test_solve, dim3((nk + 63) / 64), dim3(64), 0, st, nk, kvec, cdelta, sigma,
                     ds, dslam, v);
  return hipGetLastError();
}


}  // namespace wlc
