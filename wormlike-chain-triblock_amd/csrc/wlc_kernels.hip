// HIP kernels of the wormlike-chain MDE propagator step (gfx950, FP64).
//
// One contour step of the reference (wormlike_MDE_2dcomp_mod.f90:323-900 forward,
// :1044-1500 backward) is: BDF right-hand side -> per-harmonic 3D R2C FFT -> per-k
// G(k)^-1 apply -> C2R inverse FFT / N.  Here it is three kernels over the whole
// harmonic stack resident in HBM:
//   A  rhs_fft_xy : per (harmonic, z-plane): RHS from the 1-3 history planes and w,
//                   x transform (two real rows per complex FFT) and y transform in LDS
//   B  zsolve     : per (ky, kx tile) pencil with all harmonics in LDS: z transform,
//                   solve of G(k) q = P in the frame where k is the polar axis
//                   (tridiagonal in l per m), inverse z transform
//   C  ifft_xy    : per (harmonic, z-plane): inverse y and x (C2R) transforms, 1/N,
//                   store of q(s)
// The FFTs are LDS Stockham passes of radix 8/4/3/2 with twiddles from a table.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "wlc_kernels.h"

namespace wlc {

// ---------------------------------------------------------------------------
// orientation tables (constant memory; read with wave-uniform indices)
// ---------------------------------------------------------------------------
__constant__ double c_zsign[kMaxM];
__constant__ double c_yh[1024];
__constant__ double c_tri[kMaxM];

hipError_t upload_orientation_tables(const OrientTables& t) {
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_zsign), t.zsign, sizeof(double) * kMaxM);
  if (e != hipSuccess) return e;
  e = hipMemcpyToSymbol(HIP_SYMBOL(c_yh), t.yh, sizeof(double) * 1024);
  if (e != hipSuccess) return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_tri), t.tri_e, sizeof(double) * kMaxM);
}

// ---------------------------------------------------------------------------
// complex helpers and small DFTs
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// multiply by DIR * i  (DIR = -1: -i, +1: +i)
template <int DIR>
__device__ __forceinline__ double2 cmul_i(double2 a) {
  return DIR < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}
// twiddle from the table exp(-2 pi i t/n); the inverse uses the conjugate
template <int DIR>
__device__ __forceinline__ double2 twiddle(double2 a, double2 w) {
  const double wy = DIR < 0 ? w.y : -w.y;
  return make_double2(a.x * w.x - a.y * wy, a.x * wy + a.y * w.x);
}

template <int R, int DIR>
struct Dft;

template <int DIR>
struct Dft<2, DIR> {
  __device__ __forceinline__ static void run(double2* v) {
    const double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int DIR>
struct Dft<4, DIR> {
  __device__ __forceinline__ static void run(double2* v) {
    const double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const double2 t2 = cadd(v[1], v[3]), t3 = cmul_i<DIR>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

template <int DIR>
struct Dft<8, DIR> {
  __device__ __forceinline__ static void run(double2* v) {
    double2 e[4] = {v[0], v[2], v[4], v[6]};
    double2 o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, DIR>::run(e);
    Dft<4, DIR>::run(o);
    const double h = 0.70710678118654752440;  // sqrt(1/2)
    // W8^k = exp(DIR 2 pi i k / 8)
    const double2 o1 = make_double2(h * (o[1].x - DIR * o[1].y), h * (o[1].y + DIR * o[1].x));
    const double2 o2 = cmul_i<DIR>(o[2]);
    const double2 o3 = make_double2(h * (-o[3].x - DIR * o[3].y), h * (-o[3].y + DIR * o[3].x));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

template <int DIR>
struct Dft<3, DIR> {
  __device__ __forceinline__ static void run(double2* v) {
    const double s = 0.86602540378443864676;  // sqrt(3)/2
    const double2 a = v[0], t1 = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    const double2 t2 = make_double2(a.x - 0.5 * t1.x, a.y - 0.5 * t1.y);
    // i * DIR * s * d
    const double2 t3 = DIR < 0 ? make_double2(s * d.y, -s * d.x) : make_double2(-s * d.y, s * d.x);
    v[0] = cadd(a, t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
  }
};

// One in-place Stockham pass of radix R over nbatch transforms of length n held in
// LDS (element j of transform b at buf[b*bs + j*es]).  Every thread loads all its
// butterflies into registers before the barrier, then stores.
template <int R, int DIR, int BPT>
__device__ __forceinline__ void stockham_pass(double2* buf, int n, int Ns, int nbatch, int es,
                                              int bs, const double2* __restrict__ tw) {
  const int nb = n / R;
  const int total = nbatch * nb;
  const int tstr = n / (Ns * R);
  const int tid = threadIdx.x, nthr = blockDim.x;
  double2 v[BPT][R];
  int dsto[BPT];
#pragma unroll
  for (int c = 0; c < BPT; ++c) {
    const int idx = tid + c * nthr;
    dsto[c] = -1;
    if (idx < total) {
      const int b = idx / nb;
      const int j = idx - b * nb;
      const double2* src = buf + b * bs;
      const int k = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) v[c][r] = src[(j + r * nb) * es];
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[c][r] = twiddle<DIR>(v[c][r], tw[k * r * tstr]);
      }
      Dft<R, DIR>::run(v[c]);
      dsto[c] = b * bs + ((j - k) * R + k) * es;
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < BPT; ++c) {
    if (dsto[c] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[dsto[c] + r * Ns * es] = v[c][r];
    }
  }
  __syncthreads();
}

// full transform along one axis; MAXE bounds (nbatch * n) / blockDim.x
template <int DIR, int MAXE>
__device__ void fft_axis(double2* buf, const AxisDev& ax, int nbatch, int es, int bs) {
  int Ns = 1;
  for (int p = 0; p < ax.npass; ++p) {
    const int R = ax.radix[p];
    if (R == 8) stockham_pass<8, DIR, (MAXE + 7) / 8>(buf, ax.n, Ns, nbatch, es, bs, ax.tw);
    else if (R == 4) stockham_pass<4, DIR, (MAXE + 3) / 4>(buf, ax.n, Ns, nbatch, es, bs, ax.tw);
    else if (R == 3) stockham_pass<3, DIR, (MAXE + 2) / 3>(buf, ax.n, Ns, nbatch, es, bs, ax.tw);
    else stockham_pass<2, DIR, (MAXE + 1) / 2>(buf, ax.n, Ns, nbatch, es, bs, ax.tw);
    Ns *= R;
  }
}

// ---------------------------------------------------------------------------
// Kernel A: RHS + x (R2C) + y passes.  grid (nz, M), LDS ny*nxh complex
// ---------------------------------------------------------------------------
template <int MAXE>
__global__ void __launch_bounds__(256) k_rhs_fft_xy(Geom g, RhsArgs a) {
  extern __shared__ double2 lds[];
  const int z = blockIdx.x, i = blockIdx.y;
  const int nx = g.nx, ny = g.ny, nxh = g.nxh;
  const long long N = (long long)nx * ny * g.nz;
  const long long plane = (long long)z * nx * ny;
  const double* q1 = (a.h[0] && i < a.hM[0]) ? a.h[0] + i * N + plane : nullptr;
  const double* q2 = (a.h[1] && i < a.hM[1]) ? a.h[1] + i * N + plane : nullptr;
  const double* q3 = (a.h[2] && i < a.hM[2]) ? a.h[2] + i * N + plane : nullptr;
  const double* w = a.w + plane;
  double* pk = reinterpret_cast<double*>(lds);  // packed rows: (row 2j, row 2j+1) -> complex
  const double ds = a.ds;
  for (int e = threadIdx.x; e < nx * ny; e += blockDim.x) {
    const double v1 = q1 ? q1[e] : 0.0, v2 = q2 ? q2[e] : 0.0, v3 = q3 ? q3[e] : 0.0;
    const double t = a.a[0] * v1 + a.a[1] * v2 + a.a[2] * v3;
    const double u = a.b[0] * v1 + a.b[1] * v2 + a.b[2] * v3;
    const double p = t - (ds * w[e]) * u;
    const int y = e / nx, x = e - y * nx;
    pk[2 * ((y >> 1) * nx + x) + (y & 1)] = p;
  }
  __syncthreads();
  fft_axis<-1, MAXE>(lds, g.ax, ny / 2, 1, nx);
  // unpack Z = A + iB into A_k = (Z_k + conj Z_{n-k})/2, B_k = (Z_k - conj Z_{n-k})/(2i)
  {
    constexpr int UPT = (MAXE + 1) / 2 + 1;
    double2 ra[UPT], rb[UPT];
    int dst[UPT];
    const int tot = (ny / 2) * nxh;
#pragma unroll
    for (int c = 0; c < UPT; ++c) {
      const int idx = threadIdx.x + c * blockDim.x;
      dst[c] = -1;
      if (idx < tot) {
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
  double2* out = a.spec + ((long long)i * g.nz + z) * ny * nxh;
  for (int e = threadIdx.x; e < ny * nxh; e += blockDim.x) out[e] = lds[e];
}

// ---------------------------------------------------------------------------
// Kernel C: inverse y + x (C2R) passes, scale, store.  grid (nz, M)
// ---------------------------------------------------------------------------
template <int MAXE>
__global__ void __launch_bounds__(256) k_ifft_xy(Geom g, InvArgs a) {
  extern __shared__ double2 lds[];
  const int z = blockIdx.x, i = blockIdx.y;
  const int nx = g.nx, ny = g.ny, nxh = g.nxh;
  const double2* in = a.spec + ((long long)i * g.nz + z) * ny * nxh;
  for (int e = threadIdx.x; e < ny * nxh; e += blockDim.x) lds[e] = in[e];
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
      const int idx = threadIdx.x + c * blockDim.x;
      dst[c] = -1;
      if (idx < tot) {
        const int j = idx / nx, k = idx - j * nx;
        double2 A, B;
        if (k <= half) {
          A = lds[(2 * j) * nxh + k];
          B = lds[(2 * j + 1) * nxh + k];
          if (k == 0 || k == half) { A.y = 0.0; B.y = 0.0; }
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
  for (int e = threadIdx.x; e < nx * ny; e += blockDim.x) {
    const int y = e / nx, x = e - y * nx;
    out[e] = pk[2 * ((y >> 1) * nx + x) + (y & 1)] * sc;
  }
}

// ---------------------------------------------------------------------------
// k-aligned solve (all harmonics of one k-point in registers)
// ---------------------------------------------------------------------------
__host__ __device__ constexpr bool yh_pat(int l, int m, int mp) {
  return ((m >= 0) != (mp >= 0)) ? false
                                 : ((m >= 0) ? (((m + mp + l) & 1) == 0) : (((m + mp + l) & 1) == 1));
}
__host__ __device__ constexpr int yh_nnz(int l) {
  int c = 0;
  for (int m = -l; m <= l; ++m)
    for (int mp = -l; mp <= l; ++mp)
      if (yh_pat(l, m, mp)) ++c;
  return c;
}
__host__ __device__ constexpr int yh_off(int l) {
  int o = 0;
  for (int a = 0; a < l; ++a) o += yh_nnz(a);
  return o;
}
static_assert(yh_off(kMaxL + 1) <= 1024, "packed y-rotation table too small");

template <int L>
struct Solver {
  static constexpr int M = (L + 1) * (L + 1);

  __device__ __forceinline__ static void zr(double2 (&v)[M], double ca, double sa) {
    double cm = ca, sm = sa;
#pragma unroll
    for (int m = 1; m <= L; ++m) {
      if (m > 1) {
        const double c2 = cm * ca - sm * sa;
        sm = sm * ca + cm * sa;
        cm = c2;
      }
#pragma unroll
      for (int l = m; l <= L; ++l) {
        const int ip = l * l + l + m, in = l * l + l - m;
        const double2 vp = v[ip], vn = v[in];
        const double zp = c_zsign[ip] * sm, zn = c_zsign[in] * sm;
        v[ip] = make_double2(cm * vp.x + zp * vn.x, cm * vp.y + zp * vn.y);
        v[in] = make_double2(cm * vn.x + zn * vp.x, cm * vn.y + zn * vp.y);
      }
    }
  }

  template <bool TR>
  __device__ __forceinline__ static void yh(double2 (&v)[M]) {
#pragma unroll
    for (int l = 1; l <= L; ++l) {
      double2 o[2 * L + 1];
#pragma unroll
      for (int m = 0; m < 2 * l + 1; ++m) o[m] = make_double2(0.0, 0.0);
      int e = yh_off(l);
#pragma unroll
      for (int m = -l; m <= l; ++m)
#pragma unroll
        for (int mp = -l; mp <= l; ++mp)
          if (yh_pat(l, m, mp)) {
            const double c = c_yh[e++];
            if (!TR) {
              const double2 x = v[l * l + l + mp];
              o[m + l].x += c * x.x;
              o[m + l].y += c * x.y;
            } else {
              const double2 x = v[l * l + l + m];
              o[mp + l].x += c * x.x;
              o[mp + l].y += c * x.y;
            }
          }
#pragma unroll
      for (int m = -l; m <= l; ++m) v[l * l + l + m] = o[m + l];
    }
  }

  // tridiagonal solve per m of (diag(c + dslam l(l+1)) + i gk Rz) y = v
  __device__ __forceinline__ static void tri(double2 (&v)[M], double cdelta, double dslam, double gk) {
#pragma unroll
    for (int a = 0; a <= L; ++a) {
      double rb[L + 1];
      double beta = cdelta + dslam * a * (a + 1);
      rb[a] = 1.0 / beta;
#pragma unroll
      for (int l = a + 1; l <= L; ++l) {
        const double bb = gk * c_tri[(l - 1) * (l - 1) + (l - 1) + a];
        beta = (cdelta + dslam * l * (l + 1)) + bb * bb * rb[l - 1];
        rb[l] = 1.0 / beta;
      }
#pragma unroll
      for (int sgn = 0; sgn < (a == 0 ? 1 : 2); ++sgn) {
        const int m = sgn ? -a : a;
        double2 gg[L + 1];
        gg[a] = v[a * a + a + m];
#pragma unroll
        for (int l = a + 1; l <= L; ++l) {
          const double mu = gk * c_tri[(l - 1) * (l - 1) + (l - 1) + m] * rb[l - 1];
          const double2 vl = v[l * l + l + m];
          gg[l] = make_double2(vl.x + mu * gg[l - 1].y, vl.y - mu * gg[l - 1].x);
        }
        double2 y = make_double2(gg[L].x * rb[L], gg[L].y * rb[L]);
        v[L * L + L + m] = y;
#pragma unroll
        for (int l = L - 1; l >= a; --l) {
          const double bb = gk * c_tri[l * l + l + m];
          y = make_double2((gg[l].x + bb * y.y) * rb[l], (gg[l].y - bb * y.x) * rb[l]);
          v[l * l + l + m] = y;
        }
      }
    }
  }

  __device__ __forceinline__ static void solve(double2 (&v)[M], double kx, double ky, double kz,
                                               double cdelta, double sigma, double ds, double dslam) {
    const double kap2 = kx * kx + ky * ky + kz * kz;
    if (kap2 == 0.0) {
#pragma unroll
      for (int l = 0; l <= L; ++l) {
        const double r = 1.0 / (cdelta + dslam * l * (l + 1));
#pragma unroll
        for (int m = -l; m <= l; ++m) {
          v[l * l + l + m].x *= r;
          v[l * l + l + m].y *= r;
        }
      }
      return;
    }
    if (L == 0) {
      const double r = 1.0 / cdelta;
      v[0].x *= r;
      v[0].y *= r;
      return;
    }
    const double kap = sqrt(kap2);
    const double kp = sqrt(kx * kx + ky * ky);
    const double ct = kz / kap, st = kp / kap;
    double cp = 1.0, sp = 0.0;
    if (kp > 0.0) {
      cp = kx / kp;
      sp = ky / kp;
    }
    const double cq = -sp, sq = cp;  // phi + pi/2
    zr(v, cq, sq);
    yh<false>(v);
    zr(v, ct, st);
    yh<true>(v);
    tri(v, cdelta, dslam, sigma * ds * kap);
    yh<false>(v);
    zr(v, ct, -st);
    yh<true>(v);
    zr(v, cq, -sq);
  }
};

__device__ __forceinline__ int signed_k(int j, int n) { return (j <= n / 2) ? j : j - n; }

// ---------------------------------------------------------------------------
// Kernel B: z pass + solve + inverse z pass.  grid (ceil(nxh/TX), ny), LDS M*TX*nz
// ---------------------------------------------------------------------------
template <int L, int MAXE>
__global__ void __launch_bounds__(256) k_zsolve(Geom g, SolveArgs a) {
  constexpr int M = (L + 1) * (L + 1);
  extern __shared__ double2 lds[];
  const int TX = a.TX;
  const int kx0 = blockIdx.x * TX, ky = blockIdx.y;
  const int nz = g.nz, ny = g.ny, nxh = g.nxh;
  const int col = TX * nz;
  const long long zstride = (long long)ny * nxh;
  const long long istride = (long long)nz * zstride;
  double2* spec = a.spec + (long long)ky * nxh;
  // load [i][z][tx] (tx fastest in memory) into LDS [i][tx][z]
  for (int e = threadIdx.x; e < M * col; e += blockDim.x) {
    const int i = e / col;
    const int rem = e - i * col;
    const int z = rem / TX, tx = rem - z * TX;
    const int kx = kx0 + tx;
    double2 v = make_double2(0.0, 0.0);
    if (kx < nxh) v = spec[i * istride + z * zstride + kx];
    lds[(i * TX + tx) * nz + z] = v;
  }
  __syncthreads();
  fft_axis<-1, MAXE>(lds, g.az, M * TX, 1, nz);
  if (a.solve) {
    const int nxs = g.nx;
    const double kyv = signed_k(ky, ny) * a.kfac[1];
    for (int p = threadIdx.x; p < col; p += blockDim.x) {
      const int tx = p / nz, z = p - tx * nz;
      const int kx = kx0 + tx;
      if (kx >= nxh) continue;
      const double kxv = signed_k(kx, nxs) * a.kfac[0];
      const double kzv = signed_k(z, nz) * a.kfac[2];
      double2 v[M];
#pragma unroll
      for (int i = 0; i < M; ++i) v[i] = lds[(i * TX + tx) * nz + z];
      Solver<L>::solve(v, kxv, kyv, kzv, a.cdelta, a.sigma, a.ds, a.dslam);
#pragma unroll
      for (int i = 0; i < M; ++i) lds[(i * TX + tx) * nz + z] = v[i];
    }
    __syncthreads();
  }
  fft_axis<+1, MAXE>(lds, g.az, M * TX, 1, nz);
  for (int e = threadIdx.x; e < M * col; e += blockDim.x) {
    const int i = e / col;
    const int rem = e - i * col;
    const int z = rem / TX, tx = rem - z * TX;
    const int kx = kx0 + tx;
    if (kx < nxh) spec[i * istride + z * zstride + kx] = lds[(i * TX + tx) * nz + z];
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
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

hipError_t launch_rhs_fft(const Geom& g, const RhsArgs& a, hipStream_t st) {
  const size_t lds = sizeof(double2) * (size_t)g.ny * g.nxh;
  const long long elems = (long long)g.ny * g.nxh;
  const int me = pick_maxe(elems, 256);
  dim3 grid(g.nz, a.M), block(256);
  void (*k)(Geom, RhsArgs) = nullptr;
  switch (me) {
    case 8: k = k_rhs_fft_xy<8>; break;
    case 16: k = k_rhs_fft_xy<16>; break;
    case 32: k = k_rhs_fft_xy<32>; break;
    case 64: k = k_rhs_fft_xy<64>; break;
    default: return hipErrorInvalidValue;
  }
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, grid, block, lds, st, g, a);
  return hipGetLastError();
}

hipError_t launch_ifft_store(const Geom& g, const InvArgs& a, hipStream_t st) {
  const size_t lds = sizeof(double2) * (size_t)g.ny * g.nxh;
  const long long elems = (long long)g.ny * g.nxh;
  const int me = pick_maxe(elems, 256);
  dim3 grid(g.nz, a.M), block(256);
  void (*k)(Geom, InvArgs) = nullptr;
  switch (me) {
    case 8: k = k_ifft_xy<8>; break;
    case 16: k = k_ifft_xy<16>; break;
    case 32: k = k_ifft_xy<32>; break;
    case 64: k = k_ifft_xy<64>; break;
    default: return hipErrorInvalidValue;
  }
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, grid, block, lds, st, g, a);
  return hipGetLastError();
}

template <int L>
static void* zsolve_kernel(int me) {
  switch (me) {
    case 8: return (void*)k_zsolve<L, 8>;
    case 16: return (void*)k_zsolve<L, 16>;
    case 32: return (void*)k_zsolve<L, 32>;
    case 64: return (void*)k_zsolve<L, 64>;
  }
  return nullptr;
}

hipError_t launch_zsolve(const Geom& g, const SolveArgs& a, hipStream_t st) {
  const int M = (a.L + 1) * (a.L + 1);
  if (M != a.M) return hipErrorInvalidValue;
  const size_t lds = sizeof(double2) * (size_t)M * a.TX * g.nz;
  const int me = pick_maxe((long long)M * a.TX * g.nz, 256);
  void* k = nullptr;
  switch (a.L) {
    case 0: k = zsolve_kernel<0>(me); break;
    case 1: k = zsolve_kernel<1>(me); break;
    case 2: k = zsolve_kernel<2>(me); break;
    case 3: k = zsolve_kernel<3>(me); break;
    case 4: k = zsolve_kernel<4>(me); break;
    case 5: k = zsolve_kernel<5>(me); break;
    case 6: k = zsolve_kernel<6>(me); break;
    case 7: k = zsolve_kernel<7>(me); break;
    case 8: k = zsolve_kernel<8>(me); break;
    default: return hipErrorInvalidValue;
  }
  if (!k) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  dim3 grid((g.nxh + a.TX - 1) / a.TX, g.ny), block(256);
  Geom gg = g;
  SolveArgs aa = a;
  void* args[] = {&gg, &aa};
  return hipLaunchKernel(k, grid, block, args, lds, st);
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

template <int L>
__global__ void k_test_solve(int nk, const double* kvec, double cdelta, double sigma, double ds,
                             double dslam, double2* v) {
  constexpr int M = (L + 1) * (L + 1);
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nk) return;
  double2 x[M];
#pragma unroll
  for (int i = 0; i < M; ++i) x[i] = v[(long long)p * M + i];
  Solver<L>::solve(x, kvec[3 * p], kvec[3 * p + 1], kvec[3 * p + 2], cdelta, sigma, ds, dslam);
#pragma unroll
  for (int i = 0; i < M; ++i) v[(long long)p * M + i] = x[i];
}

hipError_t launch_test_solve(int L, int nk, const double* kvec, double cdelta, double sigma,
                             double ds, double dslam, double2* v, hipStream_t st) {
  dim3 grid((nk + 63) / 64), block(64);
  switch (L) {
#define WLC_TS(LL) \
  case LL: hipLaunchKernelGGL(k_test_solve<LL>, grid, block, 0, st, nk, kvec, cdelta, sigma, ds, dslam, v); break;
    WLC_TS(0) WLC_TS(1) WLC_TS(2) WLC_TS(3) WLC_TS(4) WLC_TS(5) WLC_TS(6) WLC_TS(7) WLC_TS(8)
#undef WLC_TS
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace wlc
