// 64 x 64 planes of the propagator step (device code, included by wlc_plane.hip).
//
// Each axis is a radix-8 x radix-8 Stockham transform.  Pass 1 of a 64-point transform
// reads element j + 8r (r = 0..7) into the registers of thread j and writes its DFT8
// output r to position 8j + r; pass 2 reads j + 8r again, twiddles by w^(jr) and leaves
// X[j + 8r] in the registers of thread j.  Each x-direction step touches only the rows
// of the calling wave, so the x-inverse -> right-hand side -> x-forward chain
// of a contour step needs no workgroup barrier.
//
// Thread roles (256 threads):
//   row role     b = t >> 3 (row pair 0..31), j = t & 7: x = j + 8r
//   column role  c = t & 31 (kx column 0..31), jy = t >> 5: y = jy + 8r
//   coalesced    wave w: row pairs 8w..8w+7 as 16-byte pieces (right-hand side, q store)
// LDS plane (2048 complex, 32 KB):
//   x layout  row pair b (rows 2b and 2b+1 as the real and imaginary parts of one
//             complex row), element x at xs(b, x)
//   y layout  half-spectrum row y, column c at ys(y, c); column 0 holds
//             X(kx = 0) + i X(kx = 32), both real along y before the forward y pass
// Row pair b owns slots [64b, 64b + 64) in both layouts, so every x-direction step is
// private to the 8 threads of one row pair (one wave) and needs only a wave barrier.
// The XOR swizzles keep the 16 lanes of a 256-byte LDS cycle on distinct banks.
#pragma once

namespace wlc {
namespace p64 {

constexpr int kPlane = 2048;  // complex slots of the plane (32 KB)
constexpr int kLds = kPlane;

__device__ __forceinline__ int xs(int b, int x) {
  return 64 * b + (x ^ ((x >> 3) & 7) ^ ((b & 1) << 3));
}
__device__ __forceinline__ int ys(int y, int c) { return 32 * y + (c ^ (((y >> 1) & 1) << 3)); }

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Twiddles exp(-2 pi i j r / 64), r = 1..7, read from the 64-entry table at each pass
// (j = t & 7 for the x passes, j = t >> 5 for the y passes); the inverse passes use
// their conjugates.
struct Tw {
  const double2* __restrict__ t;
  __device__ __forceinline__ double2 x(int r) const { return t[(threadIdx.x & 7) * r]; }
  __device__ __forceinline__ double2 y(int r) const { return t[(threadIdx.x >> 5) * r]; }
};

__device__ __forceinline__ void load_tw(Tw& w, const double2* __restrict__ tw64) { w.t = tw64; }

// Spectrum plane (i, z) of the half spectrum: a workgroup-uniform base pointer and
// 32-bit element offsets (ky stride 4 nt nz, kx-tile stride 4 nz), so the per-thread
// address arithmetic stays in 32 bits.
template <class S>
struct SpecPlane {
  S* p;
  unsigned rs, ts;
  __device__ __forceinline__ SpecPlane(S* spec, const Geom& g, int i, int z)
      : p(spec + spec_index(i, 0, 0, z, 64, g.nt, g.nz)),
        rs(unsigned(g.nt) * unsigned(g.nz) * kKxTile),
        ts(unsigned(g.nz) * kKxTile) {}
  __device__ __forceinline__ unsigned at(int ky, int kx) const {
    return unsigned(ky) * rs + unsigned(kx >> 2) * ts + unsigned(kx & 3);
  }
};

// Inverse y pass of half-spectrum plane (i, z) into the y layout (unscaled).  Columns
// kx = 0 and 32 enter as H0 + i H32 with H(ky) = (X(ky) + conj X(-ky)) / 2: the
// inverse of H is the real part of the inverse of X, i.e. exactly what a C2R transform
// keeps of those two columns.
template <class S>
__device__ __forceinline__ void y_inverse(double2* buf, const S* __restrict__ spec,
                                          const Geom& g, int i, int z, const Tw& w) {
  const int t = threadIdx.x, c = t & 31, jy = t >> 5;
  double2* hp = buf;  // packed column, one entry per ky (before the plane is written)
  const SpecPlane<S> sp(const_cast<S*>(spec), g, i, z);
  double2 v[8];
  if (c != 0) {
    const S* col = sp.p + sp.at(jy, c);
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = spec_ld(col + unsigned(8 * r) * sp.rs);
  }
  if (t < 64) {
    const int ky = t, kn = (64 - t) & 63;
    const double2 a = spec_ld(sp.p + sp.at(ky, 0)), an = spec_ld(sp.p + sp.at(kn, 0));
    const double2 b = spec_ld(sp.p + sp.at(ky, 32)), bn = spec_ld(sp.p + sp.at(kn, 32));
    const double2 h0 = make_double2(0.5 * (a.x + an.x), 0.5 * (a.y - an.y));
    const double2 h1 = make_double2(0.5 * (b.x + bn.x), 0.5 * (b.y - bn.y));
    hp[ky] = make_double2(h0.x - h1.y, h0.y + h1.x);
  }
  __syncthreads();
  if (c == 0) {
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = hp[jy + 8 * r];
  }
  __syncthreads();  // the scratch column is consumed before the plane overwrites it
  Dft<8, +1>::run(v);
#pragma unroll
  for (int r = 0; r < 8; ++r) buf[ys(8 * jy + r, c)] = v[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = buf[ys(jy + 8 * r, c)];
#pragma unroll
  for (int r = 1; r < 8; ++r) v[r] = twiddle<+1>(v[r], w.y(r));
  Dft<8, +1>::run(v);
#pragma unroll
  for (int r = 0; r < 8; ++r) buf[ys(jy + 8 * r, c)] = v[r];
  __syncthreads();
}

// Inverse x pass (C2R, one complex transform per row pair) from the y layout.  The
// Hermitian extension is formed while loading: Z_k = A_k + i B_k (A, B = rows 2b, 2b+1),
// Z_{64-k} = conj A_k + i conj B_k, Z_0 and Z_32 from the packed column.  On return the
// x layout holds row pair b (real part row 2b, imaginary part row 2b+1, unscaled).
__device__ __forceinline__ void x_inverse(double2* buf, const Tw& w) {
  const int b = threadIdx.x >> 3, j = threadIdx.x & 7;
  double2 v[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const bool lo = r < 4;                    // k = j + 8r < 32
    const int k = j + 8 * r;
    const int kk = lo ? k : 64 - k;           // 1..32 for the upper half
    const int col = kk & 31;                  // k = 0 and k = 32 -> packed column 0
    const double2 A = buf[ys(2 * b, col)], B = buf[ys(2 * b + 1, col)];
    double2 zv;
    if (lo) zv = make_double2(A.x - B.y, A.y + B.x);
    else zv = make_double2(A.x + B.y, B.x - A.y);
    if (col == 0) zv = lo ? make_double2(A.x, B.x) : make_double2(A.y, B.y);
    v[r] = zv;
  }
  wave_lds_sync();  // row pair b's y-layout slots are read before the x layout reuses them
  Dft<8, +1>::run(v);
#pragma unroll
  for (int r = 0; r < 8; ++r) buf[xs(b, 8 * j + r)] = v[r];
  wave_lds_sync();
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = buf[xs(b, j + 8 * r)];
#pragma unroll
  for (int r = 1; r < 8; ++r) v[r] = twiddle<+1>(v[r], w.x(r));
  Dft<8, +1>::run(v);
  wave_lds_sync();
#pragma unroll
  for (int r = 0; r < 8; ++r) buf[xs(b, j + 8 * r)] = v[r];
  wave_lds_sync();
}

// Forward x pass of the rows in the x layout, unpack into half-spectrum columns while
// loading the y pass, forward y pass, store of spectrum plane (i, z).
template <class S>
__device__ __forceinline__ void xy_forward(double2* buf, S* __restrict__ spec,
                                           const Geom& g, int i, int z, const Tw& w) {
  const int b = threadIdx.x >> 3, j = threadIdx.x & 7;
  double2 v[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = buf[xs(b, j + 8 * r)];
  Dft<8, -1>::run(v);
  wave_lds_sync();
#pragma unroll
  for (int r = 0; r < 8; ++r) buf[xs(b, 8 * j + r)] = v[r];
  wave_lds_sync();
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = buf[xs(b, j + 8 * r)];
#pragma unroll
  for (int r = 1; r < 8; ++r) v[r] = twiddle<-1>(v[r], w.x(r));
  Dft<8, -1>::run(v);
#pragma unroll
  for (int r = 0; r < 8; ++r) buf[xs(b, j + 8 * r)] = v[r];  // Z_k, k = j + 8r
  __syncthreads();
  // y pass 1; its loads unpack A_k = (Z_k + conj Z_{64-k}) / 2 (row 2bb) and
  // B_k = (Z_k - conj Z_{64-k}) / 2i (row 2bb + 1); column 0 packs (A_0 + i A_32) and
  // (B_0 + i B_32), the real values of kx = 0 and 32
  const int c = threadIdx.x & 31, jy = threadIdx.x >> 5;
  const int cn = c ? 64 - c : 32;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int y = jy + 8 * r;
    const int bb = y >> 1;
    const bool odd = y & 1;
    const double2 zk = buf[xs(bb, c)], zn = buf[xs(bb, cn)];
    double2 val = odd ? make_double2(0.5 * (zk.y + zn.y), -0.5 * (zk.x - zn.x))
                      : make_double2(0.5 * (zk.x + zn.x), 0.5 * (zk.y - zn.y));
    if (c == 0) val = odd ? make_double2(zk.y, zn.y) : make_double2(zk.x, zn.x);
    v[r] = val;
  }
  __syncthreads();
  Dft<8, -1>::run(v);
#pragma unroll
  for (int r = 0; r < 8; ++r) buf[ys(8 * jy + r, c)] = v[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = buf[ys(jy + 8 * r, c)];
#pragma unroll
  for (int r = 1; r < 8; ++r) v[r] = twiddle<-1>(v[r], w.y(r));
  Dft<8, -1>::run(v);
  const SpecPlane<S> sp(spec, g, i, z);
  if (c != 0) {
    S* col = sp.p + sp.at(jy, c);
#pragma unroll
    for (int r = 0; r < 8; ++r) spec_st(col + unsigned(8 * r) * sp.rs, v[r]);
  }
  __syncthreads();  // the plane buffer is free: exchange the packed column
  if (c == 0) {
#pragma unroll
    for (int r = 0; r < 8; ++r) buf[jy + 8 * r] = v[r];
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    // P = X0 + i X32 with X0, X32 Hermitian along ky: X0 = (P + conj P(-ky)) / 2,
    // X32 = (P - conj P(-ky)) / 2i
    const int ky = threadIdx.x;
    const double2 p = buf[ky], pn = buf[(64 - ky) & 63];
    spec_st(sp.p + sp.at(ky, 0), make_double2(0.5 * (p.x + pn.x), 0.5 * (p.y - pn.y)));
    spec_st(sp.p + sp.at(ky, 32), make_double2(0.5 * (p.y + pn.y), -0.5 * (p.x - pn.x)));
  }
}

// Real rows in the coalesced role: wave w owns row pairs 8w..8w+7, i.e. rows 16w..16w+15,
// a contiguous 8 KB of the plane; in step m = 0..3 lane l takes row pair
// b = 8w + 2m + (l >> 5) at x = 2 (l & 31), x + 1 (16-byte global accesses, 512 contiguous
// bytes per half wave).  q1 = newest history plane: the x layout times `scale`
// (FROM_LDS: the freshly inverted q(s), also stored to `out` when `store`) or history
// slice 0; zero unless `have1`.  With `rhs` the BDF right-hand side of the next step
// replaces the x layout.  Only the 8 KB of the calling wave is touched.
template <int ORDER, bool FROM_LDS>
__device__ __forceinline__ void rhs_rows(double2* buf, const RhsArgs& a, const Geom& g, int i,
                                         long long plane, bool have1, double scale,
                                         double* __restrict__ out, bool store, bool rhs) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long long hs = g.hs;
  const bool u1 = have1 && (FROM_LDS || i < a.hM[0]);
  const bool u2 = ORDER >= 2 && i < a.hM[1];
  const bool u3 = ORDER >= 3 && i < a.hM[2];
  // a history plane that is zero (beyond the BDF order, or a harmonic the slice lacks)
  // is read from the zero field, so every load is unconditional and the loads of all
  // four steps can be in flight together
  const double* w = a.w + plane;
  const double* h0 = (!FROM_LDS && u1) ? a.h[0] + i * hs + plane : a.zero + plane;
  const double* h1 = u2 ? a.h[1] + i * hs + plane : a.zero + plane;
  const double* h2 = u3 ? a.h[2] + i * hs + plane : a.zero + plane;
  const double2 zero = make_double2(0.0, 0.0);
  const int x = 2 * (lane & 31);
  const double a0 = a.a[0], a1 = a.a[1], a2 = a.a[2], b0 = a.b[0], b1 = a.b[1], b2 = a.b[2];
  const double ds = a.ds;
  double2 q0[4], q1[4];  // (x, x+1) of rows 2b, 2b+1 in step m
  if (FROM_LDS) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int b = 8 * wv + 2 * m + (lane >> 5);
      q0[m] = q1[m] = zero;
      if (u1) {
        const double2 z0 = buf[xs(b, x)], z1 = buf[xs(b, x + 1)];
        q0[m] = make_double2(z0.x * scale, z1.x * scale);
        q1[m] = make_double2(z0.y * scale, z1.y * scale);
      }
      if (store) {
        const int o0 = (2 * b) * 64 + x;
        *reinterpret_cast<double2*>(out + o0) = q0[m];
        *reinterpret_cast<double2*>(out + o0 + 64) = q1[m];
      }
    }
  }
  if (!rhs) {
    wave_lds_sync();
    return;
  }
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int b = 8 * wv + 2 * m + (lane >> 5);
    const int o0 = (2 * b) * 64 + x, o1 = o0 + 64;  // rows 2b, 2b+1 at x, x+1
    if (!FROM_LDS) {
      q0[m] = *reinterpret_cast<const double2*>(h0 + o0);
      q1[m] = *reinterpret_cast<const double2*>(h0 + o1);
    }
    const double2 p0 = ORDER >= 2 ? *reinterpret_cast<const double2*>(h1 + o0) : zero;
    const double2 p1 = ORDER >= 2 ? *reinterpret_cast<const double2*>(h1 + o1) : zero;
    const double2 s0 = ORDER >= 3 ? *reinterpret_cast<const double2*>(h2 + o0) : zero;
    const double2 s1 = ORDER >= 3 ? *reinterpret_cast<const double2*>(h2 + o1) : zero;
    const double2 w0 = *reinterpret_cast<const double2*>(w + o0);
    const double2 w1 = *reinterpret_cast<const double2*>(w + o1);
    buf[xs(b, x)] = make_double2(bdf_rhs(q0[m].x, p0.x, s0.x, w0.x, a0, a1, a2, b0, b1, b2, ds),
                                 bdf_rhs(q1[m].x, p1.x, s1.x, w1.x, a0, a1, a2, b0, b1, b2, ds));
    buf[xs(b, x + 1)] = make_double2(bdf_rhs(q0[m].y, p0.y, s0.y, w0.y, a0, a1, a2, b0, b1, b2, ds),
                                     bdf_rhs(q1[m].y, p1.y, s1.y, w1.y, a0, a1, a2, b0, b1, b2, ds));
  }
  wave_lds_sync();
}

// A: right-hand side from the history slices, forward x and y passes
template <int ORDER, class S>
__device__ __forceinline__ void k_rhs64_body(const Geom& g, const RhsArgs& a) {
  __shared__ double2 buf[kLds];
  int z, i;
  plane_of_block(g.nz, z, i);
  Tw w;
  load_tw(w, g.ax.tw);
  rhs_rows<ORDER, false>(buf, a, g, i, (long long)z * g.zs, true, 1.0, nullptr, false, true);
  xy_forward(buf, reinterpret_cast<S*>(a.spec), g, i, z, w);
}

template <int ORDER, class S>
__global__ void __launch_bounds__(256) k_rhs64(Geom g, RhsArgs a) {
  k_rhs64_body<ORDER, S>(g, a);
}

// the same for one instance of a batch per blockIdx.z
template <int ORDER, class S>
__global__ void __launch_bounds__(256) k_rhs64_b(const PlaneJob* __restrict__ jobs) {
  const PlaneJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_rhs64_body<ORDER, S>(j.g, j.a);
}

// C: inverse y and x passes, 1/N, store of q(s)
template <class S>
__device__ __forceinline__ void k_inv64_body(const Geom& g, const InvArgs& c) {
  __shared__ double2 buf[kLds];
  int z, i;
  plane_of_block(g.nz, z, i);
  Tw w;
  load_tw(w, g.ax.tw);
  y_inverse(buf, reinterpret_cast<const S*>(c.spec), g, i, z, w);
  x_inverse(buf, w);
  const long long plane = (long long)z * g.zs;
  RhsArgs none{};
  none.w = c.out;  // never read (no right-hand side)
  rhs_rows<1, true>(buf, none, g, i, plane, true, c.scale, c.out + i * g.hs + plane, true, false);
}

template <class S>
__global__ void __launch_bounds__(256) k_inv64(Geom g, InvArgs c) {
  k_inv64_body<S>(g, c);
}

// the same for one instance of a batch per blockIdx.z
template <class S>
__global__ void __launch_bounds__(256) k_inv64_b(const PlaneJob* __restrict__ jobs) {
  const PlaneJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_inv64_body<S>(j.g, j.c);
}

// C of step s fused with A of the next step (see k_inv_rhs)
template <int ORDER, class S>
__device__ __forceinline__ void k_inv_rhs64_body(const Geom& g, const InvArgs& c, const RhsArgs& a) {
  __shared__ double2 buf[kLds];
  int z, i;
  plane_of_block(g.nz, z, i);
  Tw w;
  load_tw(w, g.ax.tw);
  const long long plane = (long long)z * g.zs;
  const bool inv = i < c.M;
  const bool rhs = i < a.M;
  if (inv) {
    y_inverse(buf, reinterpret_cast<const S*>(c.spec), g, i, z, w);
    x_inverse(buf, w);
  }
  rhs_rows<ORDER, true>(buf, a, g, i, plane, inv, c.scale, c.out + i * g.hs + plane, inv, rhs);
  if (rhs) xy_forward(buf, reinterpret_cast<S*>(a.spec), g, i, z, w);
}

template <int ORDER, class S>
__global__ void __launch_bounds__(256) k_inv_rhs64(Geom g, InvArgs c, RhsArgs a) {
  k_inv_rhs64_body<ORDER, S>(g, c, a);
}

// the same for one instance of a batch per blockIdx.z
template <int ORDER, class S>
__global__ void __launch_bounds__(256) k_inv_rhs64_b(const PlaneJob* __restrict__ jobs) {
  const PlaneJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_inv_rhs64_body<ORDER, S>(j.g, j.c, j.a);
}

}  // namespace p64
}  // namespace wlc
