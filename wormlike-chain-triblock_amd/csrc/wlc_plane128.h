// 128 x 128 planes of the propagator step (device code, included by wlc_plane.hip).
//
// The same scheme as the 64 x 64 kernels, one plane per workgroup of 512 threads with the
// plane (64 row pairs of 128 complex values, 128 KB) in dynamic LDS, so one workgroup per
// CU.  A 128-point transform is held by 8 threads with 16 values each: pass 1 reads
// element j + 8r (r = 0..15) into thread j, runs a 16-point DFT over r, twiddles output
// k2 by W128^(j k2) and writes it to position k2 + 16 j; pass 2 reads positions
// k2 + 16 m (m = 0..7) for k2 = j and j + 8, runs two 8-point DFTs over m and writes
// X(k2 + 16 k1) back to the same positions.  Thread j then holds X(j + 8r) (r even:
// k2 = j, k1 = r / 2; r odd: k2 = j + 8, k1 = (r - 1) / 2).
//
// Thread roles (512 threads, 8 waves):
//   row role     b = t >> 3 (row pair 0..63), j = t & 7: x = j + 8r
//   column role  c = t & 63 (kx column 0..63), jy = t >> 6: y = jy + 8r
//   coalesced    wave w: row pairs 8w..8w+7, one row pair per step, 16-byte pieces
// LDS plane (8192 complex):
//   x layout  row pair b, element x at xs(b, x)
//   y layout  half-spectrum row y, column c at ys(y, c); column 0 holds
//             X(kx = 0) + i X(kx = 64)
// Row pair b owns slots [128b, 128b + 128) in both layouts (wave-private x steps).  The
// swizzle of xs keeps the 16 lanes of an LDS cycle on distinct 16-byte slots for the
// row-role accesses of both passes and for the coalesced role.
#pragma once

namespace wlc {
namespace p128 {

constexpr int kThreads = 512;
constexpr int kPlane = 8192;  // complex slots of the plane (128 KB)

__device__ __forceinline__ int xs(int b, int x) {
  return 128 * b + (x ^ ((x >> 4) & 7) ^ ((b & 1) << 3));
}
__device__ __forceinline__ int ys(int y, int c) { return 64 * y + (c ^ (((y >> 1) & 1) << 3)); }

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// position of output r of pass 2 in the two 8-point halves
__device__ __forceinline__ int p2(int r) { return (r & 1) ? 8 + (r >> 1) : (r >> 1); }

// pass 1 on the 16 values of thread j: DFT over r, twiddle W128^(j k2) (table
// exp(-2 pi i m / 128); conjugate for the inverse)
template <int DIR>
__device__ __forceinline__ void pass1(double2 (&u)[16], const double2* __restrict__ tw, int j) {
  Dft<16, DIR>::run(u);
#pragma unroll
  for (int k2 = 1; k2 < 16; ++k2) u[k2] = twiddle<DIR>(u[k2], tw[j * k2]);
}

// pass 2: u[0..7] = positions j + 16m, u[8..15] = positions j + 8 + 16m
template <int DIR>
__device__ __forceinline__ void pass2(double2 (&u)[16]) {
  Dft<8, DIR>::run(u);
  Dft<8, DIR>::run(u + 8);
}

// Spectrum plane (i, z): workgroup-uniform base pointer and 32-bit element offsets
template <class S>
struct SpecPlane {
  S* p;
  unsigned rs, ts;
  __device__ __forceinline__ SpecPlane(S* spec, const Geom& g, int i, int z)
      : p(spec + spec_index(i, 0, 0, z, 128, g.nt, g.nz)),
        rs(unsigned(g.nt) * unsigned(g.nz) * kKxTile),
        ts(unsigned(g.nz) * kKxTile) {}
  __device__ __forceinline__ unsigned at(int ky, int kx) const {
    return unsigned(ky) * rs + unsigned(kx >> 2) * ts + unsigned(kx & 3);
  }
};

// Inverse y pass of half-spectrum plane (i, z) into the y layout (unscaled).  Columns
// kx = 0 and 64 enter as H0 + i H64 with H(ky) = (X(ky) + conj X(-ky)) / 2 (what a C2R
// transform keeps of those two columns).
template <class S>
__device__ __forceinline__ void y_inverse(double2* buf, const S* __restrict__ spec,
                                          const Geom& g, int i, int z) {
  const int t = threadIdx.x, c = t & 63, jy = t >> 6;
  const double2* tw = g.ay.tw;
  double2* hp = buf;  // packed column, one entry per ky (before the plane is written)
  const SpecPlane<S> sp(const_cast<S*>(spec), g, i, z);
  double2 u[16];
  if (c != 0) {
    const S* col = sp.p + sp.at(jy, c);
#pragma unroll
    for (int r = 0; r < 16; ++r) u[r] = spec_ld(col + unsigned(8 * r) * sp.rs);
  }
  if (t < 128) {
    const int ky = t, kn = (128 - t) & 127;
    const double2 a = spec_ld(sp.p + sp.at(ky, 0)), an = spec_ld(sp.p + sp.at(kn, 0));
    const double2 b = spec_ld(sp.p + sp.at(ky, 64)), bn = spec_ld(sp.p + sp.at(kn, 64));
    const double2 h0 = make_double2(0.5 * (a.x + an.x), 0.5 * (a.y - an.y));
    const double2 h1 = make_double2(0.5 * (b.x + bn.x), 0.5 * (b.y - bn.y));
    hp[ky] = make_double2(h0.x - h1.y, h0.y + h1.x);
  }
  __syncthreads();
  if (c == 0) {
#pragma unroll
    for (int r = 0; r < 16; ++r) u[r] = hp[jy + 8 * r];
  }
  __syncthreads();  // the scratch column is consumed before the plane overwrites it
  pass1<+1>(u, tw, jy);
#pragma unroll
  for (int k2 = 0; k2 < 16; ++k2) buf[ys(k2 + 16 * jy, c)] = u[k2];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    u[m] = buf[ys(jy + 16 * m, c)];
    u[8 + m] = buf[ys(jy + 8 + 16 * m, c)];
  }
  pass2<+1>(u);
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    buf[ys(jy + 16 * m, c)] = u[m];
    buf[ys(jy + 8 + 16 * m, c)] = u[8 + m];
  }
  __syncthreads();
}

// Inverse x pass (C2R, one complex transform per row pair) from the y layout; the
// Hermitian extension is formed while loading: Z_k = A_k + i B_k (A, B = rows 2b, 2b+1),
// Z_{128-k} = conj A_k + i conj B_k, Z_0 and Z_64 from the packed column.  On return the
// x layout holds row pair b (real part row 2b, imaginary part row 2b+1, unscaled).
__device__ __forceinline__ void x_inverse(double2* buf, const Geom& g) {
  const int b = threadIdx.x >> 3, j = threadIdx.x & 7;
  double2 u[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const bool lo = r < 8;                    // k = j + 8r < 64
    const int k = j + 8 * r;
    const int kk = lo ? k : 128 - k;          // 1..64 for the upper half
    const int col = kk & 63;                  // k = 0 and k = 64 -> packed column 0
    const double2 A = buf[ys(2 * b, col)], B = buf[ys(2 * b + 1, col)];
    double2 zv;
    if (lo) zv = make_double2(A.x - B.y, A.y + B.x);
    else zv = make_double2(A.x + B.y, B.x - A.y);
    if (col == 0) zv = lo ? make_double2(A.x, B.x) : make_double2(A.y, B.y);
    u[r] = zv;
  }
  wave_lds_sync();  // row pair b's y-layout slots are read before the x layout reuses them
  pass1<+1>(u, g.ax.tw, j);
#pragma unroll
  for (int k2 = 0; k2 < 16; ++k2) buf[xs(b, k2 + 16 * j)] = u[k2];
  wave_lds_sync();
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    u[m] = buf[xs(b, j + 16 * m)];
    u[8 + m] = buf[xs(b, j + 8 + 16 * m)];
  }
  pass2<+1>(u);
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    buf[xs(b, j + 16 * m)] = u[m];
    buf[xs(b, j + 8 + 16 * m)] = u[8 + m];
  }
  wave_lds_sync();
}

// Forward x pass of the rows in the x layout, unpack into half-spectrum columns while
// loading the y pass, forward y pass, store of spectrum plane (i, z).
template <class S>
__device__ __forceinline__ void xy_forward(double2* buf, S* __restrict__ spec,
                                           const Geom& g, int i, int z) {
  const int b = threadIdx.x >> 3, j = threadIdx.x & 7;
  double2 u[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) u[r] = buf[xs(b, j + 8 * r)];
  pass1<-1>(u, g.ax.tw, j);
  wave_lds_sync();
#pragma unroll
  for (int k2 = 0; k2 < 16; ++k2) buf[xs(b, k2 + 16 * j)] = u[k2];
  wave_lds_sync();
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    u[m] = buf[xs(b, j + 16 * m)];
    u[8 + m] = buf[xs(b, j + 8 + 16 * m)];
  }
  pass2<-1>(u);
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    buf[xs(b, j + 16 * m)] = u[m];  // Z_k at natural positions
    buf[xs(b, j + 8 + 16 * m)] = u[8 + m];
  }
  __syncthreads();
  // y pass 1; its loads unpack A_k = (Z_k + conj Z_{128-k}) / 2 (row 2bb) and
  // B_k = (Z_k - conj Z_{128-k}) / 2i (row 2bb + 1); column 0 packs (A_0 + i A_64) and
  // (B_0 + i B_64), the real values of kx = 0 and 64
  const int c = threadIdx.x & 63, jy = threadIdx.x >> 6;
  const int cn = c ? 128 - c : 64;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int y = jy + 8 * r;
    const int bb = y >> 1;
    const bool odd = y & 1;
    const double2 zk = buf[xs(bb, c)], zn = buf[xs(bb, cn)];
    double2 val = odd ? make_double2(0.5 * (zk.y + zn.y), -0.5 * (zk.x - zn.x))
                      : make_double2(0.5 * (zk.x + zn.x), 0.5 * (zk.y - zn.y));
    if (c == 0) val = odd ? make_double2(zk.y, zn.y) : make_double2(zk.x, zn.x);
    u[r] = val;
  }
  __syncthreads();
  pass1<-1>(u, g.ay.tw, jy);
#pragma unroll
  for (int k2 = 0; k2 < 16; ++k2) buf[ys(k2 + 16 * jy, c)] = u[k2];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    u[m] = buf[ys(jy + 16 * m, c)];
    u[8 + m] = buf[ys(jy + 8 + 16 * m, c)];
  }
  pass2<-1>(u);
  // thread (c, jy) holds X(ky = jy + 8r) at u[p2(r)]
  const SpecPlane<S> sp(spec, g, i, z);
  if (c != 0) {
    S* col = sp.p + sp.at(jy, c);
#pragma unroll
    for (int r = 0; r < 16; ++r) spec_st(col + unsigned(8 * r) * sp.rs, u[p2(r)]);
  }
  __syncthreads();  // the plane buffer is free: exchange the packed column
  if (c == 0) {
#pragma unroll
    for (int r = 0; r < 16; ++r) buf[jy + 8 * r] = u[p2(r)];
  }
  __syncthreads();
  if (threadIdx.x < 128) {
    // P = X0 + i X64 with X0, X64 Hermitian along ky: X0 = (P + conj P(-ky)) / 2,
    // X64 = (P - conj P(-ky)) / 2i
    const int ky = threadIdx.x;
    const double2 p = buf[ky], pn = buf[(128 - ky) & 127];
    spec_st(sp.p + sp.at(ky, 0), make_double2(0.5 * (p.x + pn.x), 0.5 * (p.y - pn.y)));
    spec_st(sp.p + sp.at(ky, 64), make_double2(0.5 * (p.y + pn.y), -0.5 * (p.x - pn.x)));
  }
}

// Real rows in the coalesced role: wave w owns row pairs 8w..8w+7; in step m = 0..7 the
// wave takes row pair b = 8w + m, lane l at x = 2l, 2l + 1 of rows 2b and 2b + 1 (16-byte
// global accesses, a whole 1 KB row per instruction).  Same contract as the 64 x 64
// version: q1 = the x layout times `scale` (FROM_LDS, also stored to `out` when `store`)
// or history slice 0, zero unless `have1`; with `rhs` the BDF right-hand side of the next
// step replaces the x layout.  A zero history plane is read from the zero field.
template <int ORDER, bool FROM_LDS>
__device__ __forceinline__ void rhs_rows(double2* buf, const RhsArgs& a, const Geom& g, int i,
                                         long long plane, bool have1, double scale,
                                         double* __restrict__ out, bool store, bool rhs) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long long hs = g.hs;
  const bool u1 = have1 && (FROM_LDS || i < a.hM[0]);
  const bool u2 = ORDER >= 2 && i < a.hM[1];
  const bool u3 = ORDER >= 3 && i < a.hM[2];
  const double* w = a.w + plane;
  const double* h0 = (!FROM_LDS && u1) ? a.h[0] + i * hs + plane : a.zero + plane;
  const double* h1 = u2 ? a.h[1] + i * hs + plane : a.zero + plane;
  const double* h2 = u3 ? a.h[2] + i * hs + plane : a.zero + plane;
  const double2 zero = make_double2(0.0, 0.0);
  const int x = 2 * lane;
  const double a0 = a.a[0], a1 = a.a[1], a2 = a.a[2], b0 = a.b[0], b1 = a.b[1], b2 = a.b[2];
  const double ds = a.ds;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int b = 8 * wv + m;
    const int o0 = (2 * b) * 128 + x, o1 = o0 + 128;  // rows 2b, 2b+1 at x, x+1
    double2 q0 = zero, q1 = zero;
    if (FROM_LDS) {
      if (u1) {
        const double2 z0 = buf[xs(b, x)], z1 = buf[xs(b, x + 1)];
        q0 = make_double2(z0.x * scale, z1.x * scale);
        q1 = make_double2(z0.y * scale, z1.y * scale);
      }
      if (store) {
        *reinterpret_cast<double2*>(out + o0) = q0;
        *reinterpret_cast<double2*>(out + o1) = q1;
      }
    }
    if (rhs) {
      if (!FROM_LDS) {
        q0 = *reinterpret_cast<const double2*>(h0 + o0);
        q1 = *reinterpret_cast<const double2*>(h0 + o1);
      }
      const double2 p0 = ORDER >= 2 ? *reinterpret_cast<const double2*>(h1 + o0) : zero;
      const double2 p1 = ORDER >= 2 ? *reinterpret_cast<const double2*>(h1 + o1) : zero;
      const double2 s0 = ORDER >= 3 ? *reinterpret_cast<const double2*>(h2 + o0) : zero;
      const double2 s1 = ORDER >= 3 ? *reinterpret_cast<const double2*>(h2 + o1) : zero;
      const double2 w0 = *reinterpret_cast<const double2*>(w + o0);
      const double2 w1 = *reinterpret_cast<const double2*>(w + o1);
      buf[xs(b, x)] = make_double2(bdf_rhs(q0.x, p0.x, s0.x, w0.x, a0, a1, a2, b0, b1, b2, ds),
                                   bdf_rhs(q1.x, p1.x, s1.x, w1.x, a0, a1, a2, b0, b1, b2, ds));
      buf[xs(b, x + 1)] = make_double2(bdf_rhs(q0.y, p0.y, s0.y, w0.y, a0, a1, a2, b0, b1, b2, ds),
                                       bdf_rhs(q1.y, p1.y, s1.y, w1.y, a0, a1, a2, b0, b1, b2, ds));
    }
  }
  wave_lds_sync();
}

// A: right-hand side from the history slices, forward x and y passes
template <int ORDER, class S>
__device__ __forceinline__ void k_rhs128_body(const Geom& g, const RhsArgs& a) {
  extern __shared__ double2 buf[];
  int z, i;
  plane_of_block(g.nz, z, i);
  rhs_rows<ORDER, false>(buf, a, g, i, (long long)z * g.zs, true, 1.0, nullptr, false, true);
  xy_forward(buf, reinterpret_cast<S*>(a.spec), g, i, z);
}

template <int ORDER, class S>
__global__ void __launch_bounds__(kThreads) k_rhs128(Geom g, RhsArgs a) {
  k_rhs128_body<ORDER, S>(g, a);
}

// the same for one instance of a batch per blockIdx.z
template <int ORDER, class S>
__global__ void __launch_bounds__(kThreads) k_rhs128_b(const PlaneJob* __restrict__ jobs) {
  const PlaneJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_rhs128_body<ORDER, S>(j.g, j.a);
}

// C: inverse y and x passes, 1/N, store of q(s)
template <class S>
__device__ __forceinline__ void k_inv128_body(const Geom& g, const InvArgs& c) {
  extern __shared__ double2 buf[];
  int z, i;
  plane_of_block(g.nz, z, i);
  y_inverse(buf, reinterpret_cast<const S*>(c.spec), g, i, z);
  x_inverse(buf, g);
  const long long plane = (long long)z * g.zs;
  RhsArgs none{};
  none.w = c.out;  // never read (no right-hand side)
  rhs_rows<1, true>(buf, none, g, i, plane, true, c.scale, c.out + i * g.hs + plane, true, false);
}

template <class S>
__global__ void __launch_bounds__(kThreads) k_inv128(Geom g, InvArgs c) {
  k_inv128_body<S>(g, c);
}

// the same for one instance of a batch per blockIdx.z
template <class S>
__global__ void __launch_bounds__(kThreads) k_inv128_b(const PlaneJob* __restrict__ jobs) {
  const PlaneJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_inv128_body<S>(j.g, j.c);
}

// C of step s fused with A of the next step
template <int ORDER, class S>
__device__ __forceinline__ void k_inv_rhs128_body(const Geom& g, const InvArgs& c, const RhsArgs& a) {
  extern __shared__ double2 buf[];
  int z, i;
  plane_of_block(g.nz, z, i);
  const long long plane = (long long)z * g.zs;
  const bool inv = i < c.M;
  const bool rhs = i < a.M;
  if (inv) {
    y_inverse(buf, reinterpret_cast<const S*>(c.spec), g, i, z);
    x_inverse(buf, g);
  }
  rhs_rows<ORDER, true>(buf, a, g, i, plane, inv, c.scale, c.out + i * g.hs + plane, inv, rhs);
  if (rhs) xy_forward(buf, reinterpret_cast<S*>(a.spec), g, i, z);
}

template <int ORDER, class S>
__global__ void __launch_bounds__(kThreads) k_inv_rhs128(Geom g, InvArgs c, RhsArgs a) {
  k_inv_rhs128_body<ORDER, S>(g, c, a);
}

// the same for one instance of a batch per blockIdx.z
template <int ORDER, class S>
__global__ void __launch_bounds__(kThreads) k_inv_rhs128_b(const PlaneJob* __restrict__ jobs) {
  const PlaneJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_inv_rhs128_body<ORDER, S>(j.g, j.c, j.a);
}

}  // namespace p128
}  // namespace wlc
