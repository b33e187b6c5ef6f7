// Batched SCFT of many independent instances on one GPU (C ABI wlc_batch_*).
//
// The reference runs one SCFT per process: cal_scft loops MDE_q -> MDE_qstar -> density ->
// free_energy -> iterate (SCFT_control.f90:235-262), with the field update of AB_iterate
// (SCFT_iterate.f90:27-333).  Here every instance (its own grid, Lmax, block lengths,
// lambda, chi, box) keeps its fields, propagator store and Anderson history in device
// memory, and each contour step of the forward and backward sweeps is one launch per kernel
// variant covering every instance that selects it: the step kernels run their usual grid
// once per instance (instance = blockIdx.z), reading their arguments from a job table built
// once on the host.  The density terms, Q, totden, phi, the free energy and the simple /
// Anderson field update (FP64) are batched kernels as well, so one SCFT iteration of all
// instances is a fixed launch sequence with no host decision inside it; it is captured
// once in a HIP graph and replayed.  The instances are split into two lanes of about equal
// bytes whose sweeps run on two streams at once (launches overlap, the latency-bound
// small-grid ones among them); the observables and the field update follow the join.  A converged instance sets its skip flag and every later launch returns
// at once for it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/wlc_mde.h"
#include "wlc_host.h"
#include "wlc_internal.h"
#include "wlc_kernels.h"

namespace wlc {
namespace {

constexpr int kMaxHist = 16;      // Anderson history depth limit
constexpr int kParts = 256;       // partial sums per instance (as the handle's reductions)
constexpr int kMixParts = 128;    // partial blocks of the field-update reductions
constexpr int kQ = 32;            // reduced quantities of the force kernel

// per-instance state of the field update (device)
struct MixState {
  int iter;       // iterations done
  int done;       // converged or out of iterations; copied to the skip flag by commit
  int converged;
  int nr;         // Anderson coefficients in c
  int nvalid;     // history entries usable by the next Anderson step (reset by a restart)
  int restarts;
  double lam;     // current Anderson step (adaptive mode), else the fixed one
  double rbest;   // smallest largest-species residual of the Anderson phase so far
  double Qint, totden, lnQ, F, fe;  // Qint = V Q (the volume integral of q_0(Ns))
  double resid[3];
  double G[(kMaxHist + 1) * (kMaxHist + 1)];  // Gram matrix <dW_i, dW_j> of the history slots
  double c[kMaxHist];
};

struct MixOpts {
  int n_simple, n_hist, form, max_iter;
  double lam_s, lam_a, tol;
  int adaptive;
  double lam_max;
  int compact;  // rebuild the launch sequence over the advancing instances (host side)
};

struct DeltaJob {
  Geom g;
  double* slice;
  int M;
  const int* skip;
};
struct FillJob {
  double* p;
  long long n;
  const int* skip;
};
struct DensJob {
  Geom g;
  const double* q;
  const double* qs;
  double* rho0;
  int lim0;
  double w0;
  double* rho1;
  int lim1;
  double w1;
  const int* skip;
};
struct SumJob {
  Geom g;
  const double* src;
  int nf;
  double dV;
  double* partial;
  double* result;
  const int* skip;
};
struct PhiJob {
  Geom g;
  const double* rho;
  int nblk;
  double V;
  const double* totden;
  double* phi;
  const int* skip;
};
struct MixJob {
  Geom g;
  int nblk;
  double chi[9];
  double f[3];
  double V, dV;
  double kappa;           // half the mean pair chi: gain of the incompressibility force
  double* w;              // [nblk] padded fields
  const double* phi;
  double* wout;           // [n_hist + 1][nblk] padded fields
  double* dw;             // [n_hist + 1][nblk]
  double* partial;        // [kMixParts][kQ]
  MixState* st;
  int* skip_out;          // the instance's skip flag (written by commit)
  const int* skip;
};

__device__ __forceinline__ double simp_w(int j) { return (j & 1) ? (4.0 / 3.0) : (2.0 / 3.0); }

__global__ void kb_set_delta(const DeltaJob* __restrict__ jobs) {
  const DeltaJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  const Geom& g = j.g;
  if ((int)blockIdx.y >= g.nz * j.M) return;
  const int nxy = g.nx * g.ny;
  const int z = blockIdx.y % g.nz, i = blockIdx.y / g.nz;
  double* p = j.slice + i * g.hs + (long long)z * g.zs;
  const double v = i == 0 ? 1.0 : 0.0;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nxy; e += gridDim.x * blockDim.x) p[e] = v;
}

__global__ void kb_fill(const FillJob* __restrict__ jobs) {
  const FillJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < j.n;
       e += (long long)gridDim.x * blockDim.x)
    j.p[e] = 0.0;
}

// rho_t += w_t sum_{i < lim_t} q_i qs_i (the handle's density term, same arithmetic)
__global__ void __launch_bounds__(256) kb_density(const DensJob* __restrict__ jobs) {
  const DensJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  const Geom& g = j.g;
  if ((int)blockIdx.y >= g.nz) return;
  const int nxy2 = (g.nx * g.ny) >> 1;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nxy2) return;
  const long long r = (long long)blockIdx.y * g.zs + 2 * e;
  const long long hs = g.hs;
  const double* q = j.q;
  const double* qs = j.qs;
  double2 s = make_double2(0.0, 0.0), s0 = s, s1 = s;
  const int lim0 = j.lim0, lim1 = j.rho1 ? j.lim1 : 0;
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
  if (j.rho0) {
    double2* p = reinterpret_cast<double2*>(j.rho0 + r);
    const double2 o = *p;
    *p = make_double2(__builtin_fma(j.w0, s0.x, o.x), __builtin_fma(j.w0, s0.y, o.y));
  }
  if (j.rho1) {
    double2* p = reinterpret_cast<double2*>(j.rho1 + r);
    const double2 o = *p;
    *p = make_double2(__builtin_fma(j.w1, s1.x, o.x), __builtin_fma(j.w1, s1.y, o.y));
  }
}

// periodic-Simpson partial sums of sum_{k < nf} src_k (kParts blocks per instance) and their
// total, with the handle's reduction order
__global__ void __launch_bounds__(256) kb_wsum_partial(const SumJob* __restrict__ jobs) {
  const SumJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  const Geom& g = j.g;
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
    for (int k = 0; k < j.nf; ++k) f += j.src[k * g.hs + pr];
    acc += simp_w(x) * simp_w(y) * simp_w(z) * f;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) j.partial[blockIdx.x] = red[0] * j.dV;
}

__global__ void __launch_bounds__(256) kb_wsum_final(const SumJob* __restrict__ jobs) {
  const SumJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < kParts; i += blockDim.x) acc += j.partial[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *j.result = red[0];
}

__global__ void kb_phi(const PhiJob* __restrict__ jobs) {
  const PhiJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  const Geom& g = j.g;
  if ((int)blockIdx.y >= g.nz * j.nblk) return;
  const int nxy = g.nx * g.ny;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nxy) return;
  const long long r = (long long)(blockIdx.y % g.nz) * g.zs + (blockIdx.y / g.nz) * g.hs + e;
  j.phi[r] = j.rho[r] * (j.V / *j.totden);
}

__device__ __forceinline__ int hist_slot(int i, int nh) {
  const int m = i % (nh + 1);
  return m < 0 ? m + nh + 1 : m;
}

// Field update, pass 1 (SCFT_iterate.f90:140-333 generalised to n blocks): per point the
// force t_a = s_a - mean_a s_a + kappa (sum_b phi_b - 1) + xi - w_a with s_a = sum_{b /= a}
// chi_ab (phi_b - f_b), xi = mean field and kappa half the mean pair chi (for two blocks the
// reference's force exactly; the common mode drives incompressibility), Wout_a = w_a + t_a
// and dW_a = Wout_a - w_a into history slot k; simple mixing w_a += lam_s t_a.  Block
// partial sums: field-energy integrand, <xi>, max |t_a|, <dW_a^2>, <Wout_a^2>, <Wout_a>
// and the new row <dW_k, dW_(k - m)> of the Gram matrix (m = 0..nr).
__global__ void __launch_bounds__(256) kb_mix_force(const MixJob* __restrict__ jobs, MixOpts o) {
  const MixJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  const Geom& g = j.g;
  const int nb = j.nblk, nh = o.n_hist;
  const int it = j.st->iter + 1;
  const bool simple = it <= o.n_simple;
  const bool write_hist = it >= o.n_simple;
  const int k = simple ? 0 : hist_slot(it - o.n_simple, nh);
  const int nr = simple ? 0 : min(j.st->nvalid, nh);
  const long long hs = g.hs;
  double acc[kQ];
#pragma unroll
  for (int q = 0; q < kQ; ++q) acc[q] = 0.0;
  const long long N = (long long)g.nx * g.ny * g.nz;
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < N;
       r += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(r % g.nx);
    const long long t = r / g.nx;
    const int y = (int)(t % g.ny), z = (int)(t / g.ny);
    const long long pr = (long long)z * g.zs + (long long)y * g.nx + x;
    const double W = simp_w(x) * simp_w(y) * simp_w(z) * j.dV;
    double w[3] = {0.0, 0.0, 0.0}, ph[3] = {0.0, 0.0, 0.0};
    double xi = 0.0, sphi = 0.0;
    for (int a = 0; a < nb; ++a) {
      w[a] = j.w[a * hs + pr];
      ph[a] = j.phi[a * hs + pr];
      xi += w[a];
      sphi += ph[a];
    }
    xi /= nb;
    double e = 0.0;
    for (int a = 0; a < nb; ++a) {
      for (int b = a + 1; b < nb; ++b) e += j.chi[a * 3 + b] * ph[a] * ph[b];
      e -= w[a] * ph[a];
    }
    e -= xi * (1.0 - sphi);
    acc[0] += W * e;
    acc[1] += W * xi;
    double dwv[3] = {0.0, 0.0, 0.0};
    double sa[3] = {0.0, 0.0, 0.0}, sbar = 0.0;
    for (int a = 0; a < nb; ++a) {
      for (int b = 0; b < nb; ++b)
        if (b != a) sa[a] += j.chi[a * 3 + b] * (ph[b] - j.f[b]);
      sbar += sa[a];
    }
    sbar /= nb;
    const double inc = j.kappa * (sphi - 1.0);
    for (int a = 0; a < nb; ++a) {
      const double ta = sa[a] - sbar + inc + xi - w[a];
      const double wo = w[a] + ta;
      dwv[a] = wo - w[a];
      acc[3 + a] = fmax(acc[3 + a], fabs(ta));
      acc[6 + a] += W * dwv[a] * dwv[a];
      acc[9 + a] += W * wo * wo;
      acc[12 + a] += W * wo;
      if (write_hist) {
        j.wout[((long long)k * nb + a) * hs + pr] = wo;
        j.dw[((long long)k * nb + a) * hs + pr] = dwv[a];
      }
      if (simple) j.w[a * hs + pr] = w[a] + o.lam_s * ta;
    }
    for (int m = 1; m <= nr; ++m) {
      const int km = hist_slot(k - m, nh);
      double d = 0.0;
      for (int a = 0; a < nb; ++a) d += dwv[a] * j.dw[((long long)km * nb + a) * hs + pr];
      acc[15 + m] += W * d;
    }
  }
  // block reduction: max for 3..5, sums for the rest
  __shared__ double red[256];
  for (int q = 0; q < kQ; ++q) {
    if (q == 2 || q == 15 || q > 15 + nr) continue;  // unused slots (15: <dW_k, dW_k> = sum of 6..8)
    const bool mx = q >= 3 && q <= 5;
    red[threadIdx.x] = acc[q];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s)
        red[threadIdx.x] = mx ? fmax(red[threadIdx.x], red[threadIdx.x + s])
                              : red[threadIdx.x] + red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) j.partial[blockIdx.x * kQ + q] = red[0];
    __syncthreads();
  }
}

// LU with partial pivoting and solve of the n x n system a x = b (row-major a), the
// elimination of the host update (dgetrf + dgetrs as lapacksolver calls them,
// matrix_inverse_mod.f90:129-200); b returns x
__device__ void lu_solve(double* a, double* b, int n) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (fabs(a[i * n + k]) > fabs(a[p * n + k])) p = i;
    if (p != k) {
      for (int c = 0; c < n; ++c) {
        const double t = a[k * n + c];
        a[k * n + c] = a[p * n + c];
        a[p * n + c] = t;
      }
      const double t = b[k];
      b[k] = b[p];
      b[p] = t;
    }
    if (a[k * n + k] == 0.0) continue;
    for (int i = k + 1; i < n; ++i) {
      const double t = a[i * n + k] / a[k * n + k];
      for (int c = k; c < n; ++c) a[i * n + c] -= t * a[k * n + c];
      b[i] -= t * b[k];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    double t = b[k];
    for (int c = k + 1; c < n; ++c) t -= a[k * n + c] * b[c];
    b[k] = a[k * n + k] != 0.0 ? t / a[k * n + k] : 0.0;
  }
}

// Field update, pass 2 (one workgroup per instance): totals of pass 1 in a fixed order,
// F = -ln Q + field energy, residuals (max |t| while mixing simply; sqrt(<dW,dW>/<Wg,Wg>)
// with Wg = Wout in the gauge <xi> = (1/n) sum_a sum_{b /= a} chi_ab f_b in the Anderson
// phase), the new Gram row, the Anderson coefficients (U c = V, SCFT_iterate.f90:239-309)
// and the convergence test (SCFT_iterate.f90:1146).
__global__ void __launch_bounds__(64) kb_mix_solve(const MixJob* __restrict__ jobs, MixOpts o) {
  const MixJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  __shared__ double tot[kQ];
  const int nh = o.n_hist, nb = j.nblk;
  MixState* st = j.st;
  const int it = st->iter + 1;
  const bool simple = it <= o.n_simple;
  const int k = simple ? 0 : hist_slot(it - o.n_simple, nh);
  const int nr = simple ? 0 : min(st->nvalid, nh);
  if (threadIdx.x < kQ) {
    const int q = threadIdx.x;
    const bool mx = q >= 3 && q <= 5;
    double s = 0.0;
    if (!(q == 2 || q == 15 || q > 15 + nr))
      for (int p = 0; p < kMixParts; ++p)
        s = mx ? fmax(s, j.partial[p * kQ + q]) : s + j.partial[p * kQ + q];
    tot[q] = s;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double V = j.V;
  st->lnQ = log(st->Qint / V);
  st->fe = tot[0] / V;
  st->F = -st->lnQ + st->fe;
  double res[3] = {0.0, 0.0, 0.0};
  double ddk = 0.0;
  for (int a = 0; a < nb; ++a) ddk += tot[6 + a];
  double* G = st->G;
  const int ld = kMaxHist + 1;
  if (simple) {
    for (int a = 0; a < nb; ++a) res[a] = tot[3 + a];
    if (it == o.n_simple) {
      G[0] = ddk / V;
      st->nvalid = 1;
    }
  } else {
    double xbar = 0.0;
    for (int a = 0; a < nb; ++a)
      for (int b = 0; b < nb; ++b)
        if (b != a) xbar += j.chi[a * 3 + b] * j.f[b];
    xbar /= nb;
    const double sh = xbar - tot[1] / V;
    for (int a = 0; a < nb; ++a) {
      const double den = (tot[9 + a] + 2.0 * sh * tot[12 + a] + sh * sh * V) / V;
      res[a] = sqrt(fabs((tot[6 + a] / V) / den));
    }
    G[k * ld + k] = ddk / V;
    for (int m = 1; m <= nr; ++m) {
      const int km = hist_slot(k - m, nh);
      G[k * ld + km] = G[km * ld + k] = tot[15 + m] / V;
    }
    double U[kMaxHist * kMaxHist], Vv[kMaxHist];
    for (int m = 1; m <= nr; ++m) {
      const int km = hist_slot(k - m, nh);
      for (int q = 1; q <= nr; ++q) {
        const int kq = hist_slot(k - q, nh);
        U[(m - 1) * nr + (q - 1)] = G[k * ld + k] - G[k * ld + km] - G[k * ld + kq] + G[km * ld + kq];
      }
      Vv[m - 1] = G[k * ld + k] - G[k * ld + km];
    }
    lu_solve(U, Vv, nr);
    for (int m = 0; m < nr; ++m) st->c[m] = Vv[m];
  }
  st->nr = nr;
  if (!simple) {
    // history for the next step: one more entry, or (adaptive mode) a restart from the current
    // entry alone with half the step when the largest residual has grown tenfold past the best
    // since the last restart; the step grows 5 % per improving iteration up to lam_max
    double rmax = 0.0;
    for (int a = 0; a < nb; ++a) rmax = fmax(rmax, res[a]);
    if (o.adaptive && rmax > 10.0 * st->rbest) {
      st->nvalid = 1;
      st->restarts += 1;
      st->lam = fmax(0.5 * st->lam, 0.01);
      st->nr = 0;  // this iteration's update: a plain step from the current entry
      st->rbest = rmax;  // the best is measured again from the restart
    } else {
      st->nvalid = st->nvalid + 1;
      if (o.adaptive && rmax < st->rbest) st->lam = fmin(o.lam_max, 1.05 * st->lam);
      if (rmax < st->rbest) st->rbest = rmax;
    }
  }
  bool conv = true;
  for (int a = 0; a < 3; ++a) {
    st->resid[a] = a < nb ? res[a] : 0.0;
    if (a < nb && !(res[a] < o.tol)) conv = false;
  }
  // a non-finite Q, free energy or residual (fields grown past what the explicit field term of
  // the BDF step tolerates) ends the instance as failed
  bool bad = !isfinite(st->lnQ) || !isfinite(st->F);
  for (int a = 0; a < nb; ++a) bad = bad || !isfinite(res[a]);
  st->iter = it;
  st->converged = (conv && !bad) ? 1 : 0;
  st->done = (conv || bad || it >= o.max_iter) ? 1 : 0;
}

// Field update, pass 3 (Anderson phase): w = sum_m c_m Wout_m + lam sum_m c_m dW_m (form 0,
// SCFT_iterate.f90:311-324) or sum_m c_m w_m + lam sum_m c_m dW_m (form 1), accumulated
// in the order of the host update
__global__ void __launch_bounds__(256) kb_mix_combine(const MixJob* __restrict__ jobs, MixOpts o) {
  const MixJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  const MixState* st = j.st;
  const int it = st->iter;  // already advanced by pass 2
  if (it <= o.n_simple) return;
  const Geom& g = j.g;
  const int nb = j.nblk, nh = o.n_hist, nr = st->nr;
  const int k = hist_slot(it - o.n_simple, nh);
  const long long hs = g.hs;
  const long long N = (long long)g.nx * g.ny * g.nz;
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < N;
       r += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(r % g.nx);
    const long long t = r / g.nx;
    const int y = (int)(t % g.ny), z = (int)(t / g.ny);
    const long long pr = (long long)z * g.zs + (long long)y * g.nx + x;
    for (int a = 0; a < nb; ++a) {
      const double wk = j.wout[((long long)k * nb + a) * hs + pr];
      const double dk = j.dw[((long long)k * nb + a) * hs + pr];
      double wi = wk, di = dk;
      for (int m = 1; m <= nr; ++m) {
        const int km = hist_slot(k - m, nh);
        wi = wi + st->c[m - 1] * (j.wout[((long long)km * nb + a) * hs + pr] - wk);
        di = di + st->c[m - 1] * (j.dw[((long long)km * nb + a) * hs + pr] - dk);
      }
      if (o.form == 1) wi = wi - di;
      j.w[a * hs + pr] = wi + st->lam * di;
    }
  }
}

__global__ void kb_mix_commit(const MixJob* __restrict__ jobs) {
  const MixJob& j = jobs[blockIdx.z];
  if (threadIdx.x == 0 && j.st->done) *j.skip_out = 1;
}

// ---------------------------------------------------------------------------------------

enum OpType {
  kOpDelta, kOpFill, kOpRhs, kOpInv, kOpInvRhs, kOpZ, kOpDens, kOpSum, kOpFinal, kOpPhi,
  kOpMixForce, kOpMixSolve, kOpMixCombine, kOpMixCommit
};

struct Op {
  OpType type;
  int order = 0;
  int n = 0;           // jobs
  size_t off = 0;      // byte offset of the first job in the device job table
  PlaneJob prep{};     // representative jobs (kernel selection and grid)
  ZJob zrep{};
  dim3 grid{};
};

struct Inst {
  wlc_params p{};
  double chi[9] = {0};
  double f[3] = {0};
  Geom g{};
  long long N = 0, zs = 0, hs = 0;
  int nxh = 0, Ns = 0, Mmax = 0;
  int Mb[3] = {0, 0, 0}, start[3] = {0, 0, 0};
  double ds = 0, V = 0, dV = 0;
  std::vector<long long> off;
  std::vector<std::vector<double>> cw;
  double* q = nullptr;       // forward store
  double* qs = nullptr;      // backward 4-slot ring
  double2* spec = nullptr;
  double* w = nullptr;
  double* zero = nullptr;
  double* rho = nullptr;
  double* phi = nullptr;
  double* wout = nullptr;    // Anderson history
  double* dw = nullptr;
  double* partial = nullptr;
  double* mpart = nullptr;   // [kMixParts][kQ]
  double2* tw[3] = {nullptr, nullptr, nullptr};
  AxisPlan plan[3];
  int block_of(int s) const {
    for (int b = 0; b < p.nblk; ++b)
      if (s <= start[b] + p.Nblk[b]) return b;
    return p.nblk - 1;
  }
  double* qs_slot(int s) const { return qs + (long long)(s & 3) * Mmax * hs; }
};

}  // namespace
}  // namespace wlc

struct wlc_batch {
  int device = 0;
  hipStream_t st = nullptr;
  std::vector<wlc::Inst> in;
  wlc::MixOpts mo{};
  wlc::MixState* dstate = nullptr;
  int* dskip = nullptr;
  int* hskip = nullptr;          // pinned copies read after each iteration
  wlc::MixState* hstate = nullptr;
  char* djobs = nullptr;
  size_t jobs_bytes = 0;
  // the sweeps of the instances of each lane (independent instances: lane 1 runs on st2
  // concurrently with lane 0 on st), then on st the observables and the field update
  hipStream_t st2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::vector<wlc::Op> lane_ops[2], obs_ops, upd_ops;
  // instances in the launch sequence: all at first; when at most half of them are still
  // advancing, the sequence is rebuilt over those alone (a finished instance's launches
  // would only return at once, but each still costs its launch)
  std::vector<char> use;
  int included = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool swept = false;
  // per instance: iterations, convergence, wall time since the first iteration
  std::vector<double> done_ms;
  double elapsed_ms = 0.0;
  bool started = false;
};

namespace wlc {
namespace {

#define BCHK(call)                                                                         \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return host_fail(WLC_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));       \
  } while (0)

// byte table of jobs being built
struct JobTable {
  std::vector<char> bytes;
  template <class J>
  size_t add(const J& j) {
    const size_t a = alignof(J) > 16 ? alignof(J) : 16;
    size_t off = (bytes.size() + a - 1) / a * a;
    bytes.resize(off + sizeof(J));
    std::memcpy(bytes.data() + off, &j, sizeof(J));
    return off;
  }
  size_t align_to(size_t a) {
    size_t off = (bytes.size() + a - 1) / a * a;
    bytes.resize(off);
    return off;
  }
};

// groups of jobs with equal launch keys, in first-seen order
template <class J>
struct Grouper {
  std::vector<LaunchKey> keys;
  std::vector<std::vector<J>> jobs;
  void add(const LaunchKey& k, const J& j) {
    for (size_t i = 0; i < keys.size(); ++i)
      if (keys[i] == k) {
        jobs[i].push_back(j);
        return;
      }
    keys.push_back(k);
    jobs.push_back({j});
  }
};

template <class J>
void emit_group(JobTable& t, std::vector<Op>& ops, OpType type, int order, const std::vector<J>& js,
                dim3 grid = dim3()) {
  if (js.empty()) return;
  const size_t a = alignof(J) > 16 ? alignof(J) : 16;
  size_t off = t.align_to(a);
  for (const J& j : js) {
    // contiguous array of J: sizeof(J) is a multiple of alignof(J)
    size_t o = t.bytes.size();
    t.bytes.resize(o + sizeof(J));
    std::memcpy(t.bytes.data() + o, &j, sizeof(J));
  }
  Op op;
  op.type = type;
  op.order = order;
  op.n = (int)js.size();
  op.off = off;
  op.grid = grid;
  if constexpr (std::is_same<J, PlaneJob>::value) op.prep = js[0];
  if constexpr (std::is_same<J, ZJob>::value) op.zrep = js[0];
  ops.push_back(op);
}

RhsArgs rhs_args(const Inst& I, int dir, int s, double2* spec, int* order) {
  RhsArgs ra{};
  const int Ns = I.Ns;
  *order = dir > 0 ? std::min(s, 3) : std::min(Ns - s, 3);
  const int blk = I.block_of(s);
  for (int jj = 0; jj < 3; ++jj) {
    const int sh = s - dir * (jj + 1);
    if (jj < *order && sh >= 0 && sh <= Ns) {
      ra.h[jj] = dir > 0 ? I.q + I.off[sh] : I.qs_slot(sh);
      ra.hM[jj] = I.Mb[I.block_of(sh)];
    } else {
      ra.h[jj] = nullptr;
      ra.hM[jj] = 0;
    }
  }
  double cd;
  bdf(*order, ra.a, ra.b, &cd);
  ra.ds = I.ds;
  ra.w = I.w + (long long)blk * I.hs;
  ra.zero = I.zero;
  ra.M = I.Mb[blk];
  ra.spec = spec;
  return ra;
}

SolveArgs solve_args(const Inst& I, int dir, int s) {
  const int order = dir > 0 ? std::min(s, 3) : std::min(I.Ns - s, 3);
  const int blk = I.block_of(s);
  SolveArgs sa{};
  sa.spec = I.spec;
  sa.M = I.Mb[blk];
  sa.L = I.p.Lmax[blk];
  const ZOpts zo = zsolve_opts(I.p.nz, I.p.Lmax[blk], I.p.spectrum_fp32);
  sa.TX = zo.tx;
  double a[3], b[3];
  bdf(order, a, b, &sa.cdelta);
  sa.sigma = dir > 0 ? 1.0 : -1.0;
  sa.ds = I.ds;
  sa.dslam = I.ds * I.p.lambda[blk];
  for (int x = 0; x < 3; ++x) sa.kfac[x] = 2.0 * M_PI / I.p.box[x];
  sa.solve = 1;
  sa.generic = zo.generic;
  sa.occ = zo.occ;
  sa.split = zo.split;
  return sa;
}

// density term of contour point s into rho (the handle's density_at)
DensJob dens_job(const Inst& I, int s, const int* skip) {
  DensJob d{};
  d.g = I.g;
  d.q = I.q + I.off[s];
  d.qs = I.qs_slot(s);
  const int Ms = I.Mb[I.block_of(s)];
  int n = 0;
  for (int b = 0; b < I.p.nblk && n < 2; ++b) {
    const int jj = s - I.start[b];
    if (jj < 0 || jj > I.p.Nblk[b]) continue;
    double* rho = I.rho + (long long)b * I.hs;
    const int lim = std::min(Ms, I.Mb[b]);
    const double wt = I.cw[b][jj];
    if (n == 0) { d.rho0 = rho; d.lim0 = lim; d.w0 = wt; }
    else { d.rho1 = rho; d.lim1 = lim; d.w1 = wt; }
    ++n;
  }
  d.skip = skip;
  return d;
}

// one sweep direction of every instance: per step t the z solve of every instance with
// t <= Ns, then the fused inverse + next right-hand side (or the inverse alone at its
// last step); backward sweeps add the density term of the slice just written
void build_sweep(const wlc_batch* b, JobTable& t, std::vector<Op>& ops, int dir,
                 const std::vector<int>& idx) {
  const unsigned n = (unsigned)idx.size();
  int Tmax = 0;
  for (int i : idx) Tmax = std::max(Tmax, b->in[i].Ns);
  {  // q(0) / q*(Ns) = delta_{i0}
    std::vector<DeltaJob> dj;
    int gy = 0, gx = 0;
    for (int i : idx) {
      const Inst& I = b->in[i];
      const int s0 = dir > 0 ? 0 : I.Ns;
      DeltaJob d{};
      d.g = I.g;
      d.slice = dir > 0 ? I.q + I.off[0] : I.qs_slot(s0);
      d.M = I.Mb[I.block_of(s0)];
      d.skip = b->dskip + i;
      dj.push_back(d);
      gy = std::max(gy, I.g.nz * d.M);
      gx = std::max(gx, (I.g.nx * I.g.ny + 255) / 256);
    }
    emit_group(t, ops, kOpDelta, 0, dj, dim3(gx, gy, n));
  }
  if (dir < 0) {
    std::vector<FillJob> fj;
    long long mx = 0;
    for (int i : idx) {
      const Inst& I = b->in[i];
      fj.push_back(FillJob{I.rho, (long long)I.p.nblk * I.hs, b->dskip + i});
      mx = std::max(mx, (long long)I.p.nblk * I.hs);
    }
    emit_group(t, ops, kOpFill, 0, fj, dim3((unsigned)std::min<long long>((mx + 255) / 256, 65536), 1, n));
    std::vector<DensJob> dj;
    int gx = 0, gz = 0;
    for (int i : idx) {
      const Inst& I = b->in[i];
      dj.push_back(dens_job(I, I.Ns, b->dskip + i));
      gx = std::max(gx, (I.g.nx * I.g.ny / 2 + 255) / 256);
      gz = std::max(gz, I.g.nz);
    }
    emit_group(t, ops, kOpDens, 0, dj, dim3(gx, gz, n));
  }
  // kernel A of the first step
  {
    Grouper<PlaneJob> gr;
    std::vector<int> orders;
    for (int i : idx) {
      const Inst& I = b->in[i];
      const int s = dir > 0 ? 1 : I.Ns - 1;
      PlaneJob pj{};
      pj.g = I.g;
      int order;
      pj.a = rhs_args(I, dir, s, I.spec, &order);
      pj.skip = b->dskip + i;
      LaunchKey k = plane_key(I.g, 0, order, pj.a.M);
      k.gy = (unsigned)order;  // the order is part of the kernel; keep it in the key
      gr.add(k, pj);
    }
    for (size_t gi = 0; gi < gr.keys.size(); ++gi) emit_group(t, ops, kOpRhs, (int)gr.keys[gi].gy, gr.jobs[gi]);
  }
  for (int step = 1; step <= Tmax; ++step) {
    Grouper<ZJob> gz;
    for (int i : idx) {
      const Inst& I = b->in[i];
      if (step > I.Ns) continue;
      const int s = dir > 0 ? step : I.Ns - step;
      ZJob zj{};
      zj.g = I.g;
      zj.a = solve_args(I, dir, s);
      zj.skip = b->dskip + i;
      gz.add(zsolve_key(I.g, zj.a), zj);
    }
    for (size_t gi = 0; gi < gz.keys.size(); ++gi) emit_group(t, ops, kOpZ, 0, gz.jobs[gi]);
    Grouper<PlaneJob> gc, gca;
    std::vector<DensJob> dj;
    int dgx = 0, dgz = 0;
    for (int i : idx) {
      const Inst& I = b->in[i];
      if (step > I.Ns) continue;
      const int s = dir > 0 ? step : I.Ns - step;
      PlaneJob pj{};
      pj.g = I.g;
      pj.c.spec = I.spec;
      pj.c.M = I.Mb[I.block_of(s)];
      pj.c.scale = 1.0 / (double)I.N;
      pj.c.out = dir > 0 ? I.q + I.off[s] : I.qs_slot(s);
      pj.skip = b->dskip + i;
      if (step < I.Ns) {
        int order;
        pj.a = rhs_args(I, dir, s + dir, I.spec, &order);
        LaunchKey k = plane_key(I.g, 2, order, std::max(pj.c.M, pj.a.M));
        k.gy = (unsigned)order;
        gca.add(k, pj);
      } else {
        gc.add(plane_key(I.g, 1, 0, pj.c.M), pj);
      }
      if (dir < 0) {
        dj.push_back(dens_job(I, s, b->dskip + i));
        dgx = std::max(dgx, (I.g.nx * I.g.ny / 2 + 255) / 256);
        dgz = std::max(dgz, I.g.nz);
      }
    }
    for (size_t gi = 0; gi < gca.keys.size(); ++gi) emit_group(t, ops, kOpInvRhs, (int)gca.keys[gi].gy, gca.jobs[gi]);
    for (size_t gi = 0; gi < gc.keys.size(); ++gi) emit_group(t, ops, kOpInv, 0, gc.jobs[gi]);
    if (!dj.empty()) emit_group(t, ops, kOpDens, 0, dj, dim3(dgx, dgz, (unsigned)dj.size()));
  }
}

// Q, totden and phi of every instance (the handle's density_common)
void build_observables(const wlc_batch* b, JobTable& t, std::vector<Op>& ops) {
  std::vector<int> idx;
  for (int i = 0; i < (int)b->in.size(); ++i)
    if (b->use[i]) idx.push_back(i);
  const unsigned n = (unsigned)idx.size();
  std::vector<SumJob> sq, sr;
  std::vector<PhiJob> pj;
  int gx = 0, gy = 0;
  for (int i : idx) {
    const Inst& I = b->in[i];
    MixState* st = b->dstate + i;
    SumJob a{};
    a.g = I.g;
    a.src = I.q + I.off[I.Ns];
    a.nf = 1;
    a.dV = I.dV;
    a.partial = I.partial;
    a.result = &st->Qint;
    a.skip = b->dskip + i;
    sq.push_back(a);
    SumJob c = a;
    c.src = I.rho;
    c.nf = I.p.nblk;
    c.partial = I.partial + kParts;
    c.result = &st->totden;
    sr.push_back(c);
    PhiJob p{};
    p.g = I.g;
    p.rho = I.rho;
    p.nblk = I.p.nblk;
    p.V = I.V;
    p.totden = &st->totden;
    p.phi = I.phi;
    p.skip = b->dskip + i;
    pj.push_back(p);
    gx = std::max(gx, (I.g.nx * I.g.ny + 255) / 256);
    gy = std::max(gy, I.g.nz * I.p.nblk);
  }
  emit_group(t, ops, kOpSum, 0, sq, dim3(kParts, 1, n));
  emit_group(t, ops, kOpFinal, 0, sq, dim3(1, 1, n));
  emit_group(t, ops, kOpSum, 0, sr, dim3(kParts, 1, n));
  emit_group(t, ops, kOpFinal, 0, sr, dim3(1, 1, n));
  emit_group(t, ops, kOpPhi, 0, pj, dim3(gx, gy, n));
}

void build_mix(const wlc_batch* b, JobTable& t, std::vector<Op>& ops) {
  std::vector<int> idx;
  for (int i = 0; i < (int)b->in.size(); ++i)
    if (b->use[i]) idx.push_back(i);
  const unsigned n = (unsigned)idx.size();
  std::vector<MixJob> mj;
  for (int i : idx) {
    const Inst& I = b->in[i];
    MixJob m{};
    m.g = I.g;
    m.nblk = I.p.nblk;
    for (int k = 0; k < 9; ++k) m.chi[k] = I.chi[k];
    for (int k = 0; k < 3; ++k) m.f[k] = I.f[k];
    m.V = I.V;
    m.dV = I.dV;
    {
      const int nb = I.p.nblk;
      double k = 0.0;
      for (int a = 0; a < nb; ++a)
        for (int c = a + 1; c < nb; ++c) k += I.chi[a * 3 + c];
      m.kappa = nb > 1 ? k / (nb * (nb - 1)) : 0.0;
    }
    m.w = I.w;
    m.phi = I.phi;
    m.wout = I.wout;
    m.dw = I.dw;
    m.partial = I.mpart;
    m.st = b->dstate + i;
    m.skip_out = b->dskip + i;
    m.skip = b->dskip + i;
    mj.push_back(m);
  }
  emit_group(t, ops, kOpMixForce, 0, mj, dim3(kMixParts, 1, n));
  emit_group(t, ops, kOpMixSolve, 0, mj, dim3(1, 1, n));
  emit_group(t, ops, kOpMixCombine, 0, mj, dim3(kMixParts, 1, n));
  emit_group(t, ops, kOpMixCommit, 0, mj, dim3(1, 1, n));
}

int run_ops(wlc_batch* b, const std::vector<Op>& ops, hipStream_t st) {
  for (const Op& op : ops) {
    const char* base = b->djobs + op.off;
    switch (op.type) {
      case kOpDelta:
        hipLaunchKernelGGL(kb_set_delta, op.grid, dim3(256), 0, st, reinterpret_cast<const DeltaJob*>(base));
        break;
      case kOpFill:
        hipLaunchKernelGGL(kb_fill, op.grid, dim3(256), 0, st, reinterpret_cast<const FillJob*>(base));
        break;
      case kOpRhs:
        BCHK(launch_rhs_fft_batch(op.prep, op.order, reinterpret_cast<const PlaneJob*>(base), op.n, st));
        break;
      case kOpInv:
        BCHK(launch_ifft_store_batch(op.prep, reinterpret_cast<const PlaneJob*>(base), op.n, st));
        break;
      case kOpInvRhs:
        BCHK(launch_inv_rhs_batch(op.prep, op.order, reinterpret_cast<const PlaneJob*>(base), op.n, st));
        break;
      case kOpZ:
        BCHK(launch_zsolve_batch(op.zrep, reinterpret_cast<const ZJob*>(base), op.n, st));
        break;
      case kOpDens:
        hipLaunchKernelGGL(kb_density, op.grid, dim3(256), 0, st, reinterpret_cast<const DensJob*>(base));
        break;
      case kOpSum:
        hipLaunchKernelGGL(kb_wsum_partial, op.grid, dim3(256), 0, st, reinterpret_cast<const SumJob*>(base));
        break;
      case kOpFinal:
        hipLaunchKernelGGL(kb_wsum_final, op.grid, dim3(256), 0, st, reinterpret_cast<const SumJob*>(base));
        break;
      case kOpPhi:
        hipLaunchKernelGGL(kb_phi, op.grid, dim3(256), 0, st, reinterpret_cast<const PhiJob*>(base));
        break;
      case kOpMixForce:
        hipLaunchKernelGGL(kb_mix_force, op.grid, dim3(256), 0, st, reinterpret_cast<const MixJob*>(base), b->mo);
        break;
      case kOpMixSolve:
        hipLaunchKernelGGL(kb_mix_solve, op.grid, dim3(64), 0, st, reinterpret_cast<const MixJob*>(base), b->mo);
        break;
      case kOpMixCombine:
        hipLaunchKernelGGL(kb_mix_combine, op.grid, dim3(256), 0, st, reinterpret_cast<const MixJob*>(base), b->mo);
        break;
      case kOpMixCommit:
        hipLaunchKernelGGL(kb_mix_commit, op.grid, dim3(64), 0, st, reinterpret_cast<const MixJob*>(base));
        break;
    }
    BCHK(hipGetLastError());
  }
  return WLC_OK;
}

// one sweep pair of every included instance (lane 1 on st2 beside lane 0 on st), then the
// observables and, with_update, the field update on st; st has joined st2 when it returns
// (also under stream capture: the fork and join become graph edges)
int run_iteration(wlc_batch* b, bool with_update) {
  const bool two = !b->lane_ops[1].empty();
  if (two) {
    BCHK(hipEventRecord(b->ev_fork, b->st));
    BCHK(hipStreamWaitEvent(b->st2, b->ev_fork, 0));
    if (int rc = run_ops(b, b->lane_ops[1], b->st2)) return rc;
    BCHK(hipEventRecord(b->ev_join, b->st2));
  }
  if (int rc = run_ops(b, b->lane_ops[0], b->st)) return rc;
  if (two) BCHK(hipStreamWaitEvent(b->st, b->ev_join, 0));
  if (int rc = run_ops(b, b->obs_ops, b->st)) return rc;
  if (with_update) return run_ops(b, b->upd_ops, b->st);
  return WLC_OK;
}

// the included instances (b->use) split into two lanes of about equal bytes per sweep pair
// (Ns M N per instance): largest first, each to the lighter lane; each lane in index order.
// Measurement switch WLC_BATCH_LANES: 1 = all in lane 0; 2 = instances of one grid and Lmax
// kept together (one launch per kernel variant and step instead of one per lane); otherwise
// instance by instance (measured 1-2 % faster than whole groups on the 48-instance draw).
void assign_lanes(const wlc_batch* b, std::vector<int> lanes[2]) {
  std::vector<int> idx;
  for (int i = 0; i < (int)b->in.size(); ++i)
    if (b->use[i]) idx.push_back(i);
  const char* env = std::getenv("WLC_BATCH_LANES");
  const int mode = env ? std::atoi(env) : 0;
  if (mode == 1) {
    lanes[0] = idx;
    return;
  }
  auto same_group = [&](int x, int y) {
    if (mode != 2) return false;
    const wlc_params& p = b->in[x].p;
    const wlc_params& q = b->in[y].p;
    if (p.nx != q.nx || p.ny != q.ny || p.nz != q.nz || p.nblk != q.nblk ||
        p.spectrum_fp32 != q.spectrum_fp32)
      return false;
    for (int k = 0; k < p.nblk; ++k)
      if (p.Lmax[k] != q.Lmax[k]) return false;
    return true;
  };
  std::vector<std::vector<int>> groups;
  std::vector<double> cost;
  for (int i : idx) {
    const Inst& I = b->in[i];
    const double c = (double)I.Ns * (double)I.Mmax * (double)I.N;
    size_t g = 0;
    while (g < groups.size() && !same_group(groups[g][0], i)) ++g;
    if (g == groups.size()) {
      groups.emplace_back();
      cost.push_back(0.0);
    }
    groups[g].push_back(i);
    cost[g] += c;
  }
  std::vector<size_t> order(groups.size());
  for (size_t g = 0; g < order.size(); ++g) order[g] = g;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
  double load[2] = {0.0, 0.0};
  for (size_t g : order) {
    const int l = load[1] < load[0] ? 1 : 0;
    lanes[l].insert(lanes[l].end(), groups[g].begin(), groups[g].end());
    load[l] += cost[g];
  }
  for (int l = 0; l < 2; ++l) std::sort(lanes[l].begin(), lanes[l].end());
}

// the launch sequences over the included instances (b->use): the sweeps of each lane, the
// observables and the field update; the job table is uploaded again and the captured graph
// dropped (recaptured at the next iteration). Called with no work in flight on the batch's
// streams.
int rebuild_ops(wlc_batch* b) {
  JobTable t;
  std::vector<int> lanes[2];
  assign_lanes(b, lanes);
  std::vector<Op> lo[2], obs, upd;
  for (int l = 0; l < 2; ++l) {
    if (lanes[l].empty()) continue;
    build_sweep(b, t, lo[l], +1, lanes[l]);
    build_sweep(b, t, lo[l], -1, lanes[l]);
  }
  build_observables(b, t, obs);
  build_mix(b, t, upd);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  if (b->djobs) (void)hipFree(b->djobs);
  b->djobs = nullptr;
  b->jobs_bytes = t.bytes.size();
  if (hipMalloc((void**)&b->djobs, std::max<size_t>(b->jobs_bytes, 16)) != hipSuccess ||
      hipMemcpy(b->djobs, t.bytes.data(), b->jobs_bytes, hipMemcpyHostToDevice) != hipSuccess)
    return host_fail(WLC_ENOMEM, "job table upload failed");
  for (int l = 0; l < 2; ++l) b->lane_ops[l].swap(lo[l]);
  b->obs_ops.swap(obs);
  b->upd_ops.swap(upd);
  b->included = 0;
  for (char u : b->use) b->included += u ? 1 : 0;
  return WLC_OK;
}

void free_batch(wlc_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->st) (void)hipStreamSynchronize(b->st);
  if (b->st2) (void)hipStreamSynchronize(b->st2);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (Inst& I : b->in) {
    (void)hipFree(I.q);
    (void)hipFree(I.qs);
    (void)hipFree(I.spec);
    (void)hipFree(I.w);
    (void)hipFree(I.zero);
    (void)hipFree(I.rho);
    (void)hipFree(I.phi);
    (void)hipFree(I.wout);
    (void)hipFree(I.dw);
    (void)hipFree(I.partial);
    (void)hipFree(I.mpart);
    for (auto t : I.tw) (void)hipFree(t);
  }
  (void)hipFree(b->dstate);
  (void)hipFree(b->dskip);
  (void)hipFree(b->djobs);
  if (b->hskip) (void)hipHostFree(b->hskip);
  if (b->hstate) (void)hipHostFree(b->hstate);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_join) (void)hipEventDestroy(b->ev_join);
  if (b->st2) (void)hipStreamDestroy(b->st2);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
}

int check_inst(wlc_batch* b, int i) {
  if (!b) return host_fail(WLC_EINVAL, "NULL batch");
  if (i < 0 || i >= (int)b->in.size()) return host_fail(WLC_EINVAL, "instance index out of range");
  return WLC_OK;
}

hipError_t copy_fields(const Inst& I, double* dst, const double* src, int nf, bool to_device,
                       hipStream_t st) {
  const size_t row = sizeof(double) * (size_t)I.p.nx * I.p.ny;
  for (int f = 0; f < nf; ++f) {
    hipError_t e = to_device
        ? hipMemcpy2DAsync(dst + (long long)f * I.hs, sizeof(double) * I.zs, src + (long long)f * I.N,
                           row, row, I.p.nz, hipMemcpyHostToDevice, st)
        : hipMemcpy2DAsync(dst + (long long)f * I.N, row, src + (long long)f * I.hs,
                           sizeof(double) * I.zs, row, I.p.nz, hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace
}  // namespace wlc

using namespace wlc;

extern "C" {

int wlc_batch_create(const wlc_params* params, const double* chi, int n, const wlc_scft_opts* opts,
                     wlc_batch** out) {
  if (!out) return host_fail(WLC_EINVAL, "out is NULL");
  *out = nullptr;
  if (!params || n < 1) return host_fail(WLC_EINVAL, "need n >= 1 instances");
  if (!chi) return host_fail(WLC_EINVAL, "chi is NULL");
  wlc_scft_opts o{};
  if (opts) o = *opts;
  else {
    o.max_iter = 500;
    o.n_simple = 20;
    o.n_hist = 10;
    o.form = 0;
    o.lam_simple = 0.05;
    o.lam_anderson = 0.1;
    o.tol = 1e-6;
  }
  if (o.n_hist < 1 || o.n_hist > kMaxHist) return host_fail(WLC_EINVAL, "n_hist must be 1..16");
  if (o.n_simple < 0 || o.max_iter < 1 || (o.form != 0 && o.form != 1))
    return host_fail(WLC_EINVAL, "bad mixing options");
  if (o.adaptive != 0 && o.adaptive != 1) return host_fail(WLC_EINVAL, "adaptive must be 0 or 1");
  if (o.no_compact != 0 && o.no_compact != 1) return host_fail(WLC_EINVAL, "no_compact must be 0 or 1");
  for (int r : o.reserved)
    if (r) return host_fail(WLC_EINVAL, "reserved option fields must be zero");
  for (int i = 0; i < n; ++i) {
    int rc = validate_params(params + i);
    if (rc) return host_fail(rc, "instance " + std::to_string(i) + ": " + wlc_mde_last_error());
    if (params[i].keep_qstar || params[i].ckpt_every)
      return host_fail(WLC_EINVAL, "batched instances keep the full forward store (keep_qstar = ckpt_every = 0)");
  }
  wlc_batch* b = new wlc_batch();
  if (hipGetDevice(&b->device) != hipSuccess) {
    delete b;
    return host_fail(WLC_EHIP, "no HIP device");
  }
  int rc = orientation_tables_ready(b->device);
  if (rc) {
    delete b;
    return rc;
  }
  b->mo = MixOpts{o.n_simple, o.n_hist,   o.form, o.max_iter, o.lam_simple, o.lam_anderson,
                  o.tol,      o.adaptive, o.lam_max > 0.0 ? o.lam_max : 1.0, o.no_compact ? 0 : 1};
  if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->st2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess) {
    free_batch(b);
    return host_fail(WLC_EHIP, "stream creation failed");
  }
#define BALLOC(ptr, bytes)                                                                 \
  do {                                                                                     \
    if (hipMalloc((void**)&(ptr), (bytes)) != hipSuccess) {                                \
      free_batch(b);                                                                       \
      return host_fail(WLC_ENOMEM, std::string("device allocation failed: ") + #ptr);      \
    }                                                                                      \
  } while (0)
  b->in.resize(n);
  size_t need = 0;
  for (int i = 0; i < n; ++i) {
    Inst& I = b->in[i];
    const wlc_params& p = params[i];
    I.p = p;
    for (int k = 0; k < 9; ++k) I.chi[k] = chi[9 * i + k];
    I.N = (long long)p.nx * p.ny * p.nz;
    I.zs = (long long)p.nx * p.ny + 16;
    I.hs = (long long)p.nz * I.zs + 32;
    I.nxh = p.nx / 2 + 1;
    I.Ns = p.Ns;
    I.ds = 1.0 / p.Ns;
    I.V = p.box[0] * p.box[1] * p.box[2];
    I.dV = (p.box[0] / p.nx) * (p.box[1] / p.ny) * (p.box[2] / p.nz);
    int s0 = 0;
    for (int bk = 0; bk < p.nblk; ++bk) {
      I.Mb[bk] = (p.Lmax[bk] + 1) * (p.Lmax[bk] + 1);
      I.start[bk] = s0;
      s0 += p.Nblk[bk];
      I.Mmax = std::max(I.Mmax, I.Mb[bk]);
      I.f[bk] = (double)p.Nblk[bk] / p.Ns;
      std::vector<double> cw;
      simpson_weights(p.Nblk[bk], I.ds, &cw);
      I.cw.push_back(cw);
    }
    I.off.resize(I.Ns + 2);
    I.off[0] = 0;
    for (int s = 0; s <= I.Ns; ++s) I.off[s + 1] = I.off[s] + (long long)I.Mb[I.block_of(s)] * I.hs;
    const int edges[3] = {p.nx, p.ny, p.nz};
    for (int a = 0; a < 3; ++a) {
      std::string err = make_axis_plan(edges[a], &I.plan[a]);
      if (!err.empty() || I.plan[a].npass > 8) {
        free_batch(b);
        return host_fail(WLC_EINVAL, err.empty() ? "too many FFT passes" : err);
      }
    }
    const size_t fld = sizeof(double) * (size_t)p.nblk * I.hs;
    need += sizeof(double) * (size_t)I.off[I.Ns + 1] + sizeof(double) * 4 * (size_t)I.Mmax * I.hs +
            sizeof(double2) * (size_t)I.Mmax * p.ny * kx_tiles(I.nxh) * kKxTile * p.nz + 3 * fld +
            2 * (size_t)(o.n_hist + 1) * fld;
  }
  {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && need + (size_t(1) << 30) > free_b) {
      free_batch(b);
      return host_fail(WLC_ENOMEM, "the instances' stores and histories (" + std::to_string(need >> 20) +
                                       " MiB) do not fit in free device memory");
    }
  }
  for (int i = 0; i < n; ++i) {
    Inst& I = b->in[i];
    const wlc_params& p = I.p;
    const size_t fld = sizeof(double) * (size_t)p.nblk * I.hs;
    BALLOC(I.q, sizeof(double) * (size_t)I.off[I.Ns + 1]);
    BALLOC(I.qs, sizeof(double) * 4 * (size_t)I.Mmax * I.hs);
    BALLOC(I.spec, sizeof(double2) * (size_t)I.Mmax * p.ny * kx_tiles(I.nxh) * kKxTile * p.nz);
    BALLOC(I.w, fld);
    BALLOC(I.zero, sizeof(double) * (size_t)I.hs);
    BALLOC(I.rho, fld);
    BALLOC(I.phi, fld);
    BALLOC(I.wout, (size_t)(o.n_hist + 1) * fld);
    BALLOC(I.dw, (size_t)(o.n_hist + 1) * fld);
    BALLOC(I.partial, sizeof(double) * 2 * kParts);
    BALLOC(I.mpart, sizeof(double) * kMixParts * kQ);
    if (hipMemset(I.zero, 0, sizeof(double) * (size_t)I.hs) != hipSuccess ||
        hipMemset(I.w, 0, fld) != hipSuccess || hipMemset(I.phi, 0, fld) != hipSuccess ||
        hipMemset(I.wout, 0, (size_t)(o.n_hist + 1) * fld) != hipSuccess ||
        hipMemset(I.dw, 0, (size_t)(o.n_hist + 1) * fld) != hipSuccess ||
        hipMemset(I.mpart, 0, sizeof(double) * kMixParts * kQ) != hipSuccess) {
      free_batch(b);
      return host_fail(WLC_EHIP, "initialisation failed");
    }
    const int edges[3] = {p.nx, p.ny, p.nz};
    for (int a = 0; a < 3; ++a) {
      BALLOC(I.tw[a], sizeof(double2) * (size_t)edges[a]);
      if (hipMemcpy(I.tw[a], I.plan[a].tw.data(), sizeof(double2) * edges[a], hipMemcpyHostToDevice) != hipSuccess) {
        free_batch(b);
        return host_fail(WLC_EHIP, "twiddle upload failed");
      }
    }
    Geom& g = I.g;
    g.nx = p.nx; g.ny = p.ny; g.nz = p.nz; g.nxh = I.nxh;
    g.nt = kx_tiles(I.nxh);
    g.spec32 = p.spectrum_fp32;
    g.zs = I.zs;
    g.hs = I.hs;
    AxisDev* ax[3] = {&g.ax, &g.ay, &g.az};
    for (int a = 0; a < 3; ++a) {
      ax[a]->n = I.plan[a].n;
      ax[a]->npass = I.plan[a].npass;
      for (int r = 0; r < 8; ++r) ax[a]->radix[r] = r < I.plan[a].npass ? I.plan[a].radix[r] : 0;
      ax[a]->tw = I.tw[a];
    }
  }
  BALLOC(b->dstate, sizeof(MixState) * n);
  BALLOC(b->dskip, sizeof(int) * n);
#undef BALLOC
  std::vector<MixState> init(n);
  std::memset(init.data(), 0, sizeof(MixState) * n);
  for (MixState& m : init) {
    m.lam = o.lam_anderson;
    m.rbest = 1e300;
  }
  if (hipMemcpy(b->dstate, init.data(), sizeof(MixState) * n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(b->dskip, 0, sizeof(int) * n) != hipSuccess ||
      hipHostMalloc((void**)&b->hskip, sizeof(int) * n) != hipSuccess ||
      hipHostMalloc((void**)&b->hstate, sizeof(MixState) * n) != hipSuccess) {
    free_batch(b);
    return host_fail(WLC_EHIP, "state initialisation failed");
  }
  b->use.assign(n, 1);
  if (int rc = rebuild_ops(b)) {
    free_batch(b);
    return rc;
  }
  b->done_ms.assign(n, -1.0);
  *out = b;
  return WLC_OK;
}

void wlc_batch_destroy(wlc_batch* b) { free_batch(b); }

int wlc_batch_size(const wlc_batch* b) { return b ? (int)b->in.size() : WLC_EINVAL; }

int wlc_batch_launches(const wlc_batch* b, int with_update) {
  if (!b) return WLC_EINVAL;
  const size_t n = b->lane_ops[0].size() + b->lane_ops[1].size() + b->obs_ops.size();
  return (int)(with_update ? n + b->upd_ops.size() : n);
}

int wlc_batch_set_w(wlc_batch* b, int i, const double* w_host) {
  int rc = check_inst(b, i);
  if (rc) return rc;
  if (!w_host) return host_fail(WLC_EINVAL, "NULL argument");
  BCHK(hipSetDevice(b->device));
  const Inst& I = b->in[i];
  BCHK(copy_fields(I, I.w, w_host, I.p.nblk, true, b->st));
  BCHK(hipStreamSynchronize(b->st));
  b->swept = false;
  return WLC_OK;
}

int wlc_batch_get_w(wlc_batch* b, int i, double* w_host) {
  int rc = check_inst(b, i);
  if (rc) return rc;
  if (!w_host) return host_fail(WLC_EINVAL, "NULL argument");
  BCHK(hipSetDevice(b->device));
  const Inst& I = b->in[i];
  BCHK(copy_fields(I, w_host, I.w, I.p.nblk, false, b->st));
  BCHK(hipStreamSynchronize(b->st));
  return WLC_OK;
}

int wlc_batch_sweep(wlc_batch* b) {
  if (!b) return host_fail(WLC_EINVAL, "NULL batch");
  BCHK(hipSetDevice(b->device));
  int rc = run_iteration(b, false);
  if (rc) return rc;
  BCHK(hipStreamSynchronize(b->st));
  b->swept = true;
  return WLC_OK;
}

int wlc_batch_density(wlc_batch* b, int i, double* phi_host, double* Q, double* totden) {
  int rc = check_inst(b, i);
  if (rc) return rc;
  BCHK(hipSetDevice(b->device));
  BCHK(hipStreamSynchronize(b->st));
  const Inst& I = b->in[i];
  MixState s;
  BCHK(hipMemcpy(&s, b->dstate + i, sizeof(MixState), hipMemcpyDeviceToHost));
  if (Q) *Q = s.Qint / I.V;
  if (totden) *totden = s.totden;
  if (phi_host) {
    BCHK(copy_fields(I, phi_host, I.phi, I.p.nblk, false, b->st));
    BCHK(hipStreamSynchronize(b->st));
  }
  return WLC_OK;
}

int wlc_batch_get_q(wlc_batch* b, int i, int s, double* q_host) {
  int rc = check_inst(b, i);
  if (rc) return rc;
  const Inst& I = b->in[i];
  if (!q_host || s < 0 || s > I.Ns) return host_fail(WLC_EINVAL, "bad slice");
  BCHK(hipSetDevice(b->device));
  BCHK(hipStreamSynchronize(b->st));
  BCHK(copy_fields(I, q_host, I.q + I.off[s], I.Mb[I.block_of(s)], false, b->st));
  BCHK(hipStreamSynchronize(b->st));
  return WLC_OK;
}

int wlc_batch_iterate(wlc_batch* b, int niter, int use_graph, int* n_active) {
  if (!b || niter < 0) return host_fail(WLC_EINVAL, "bad arguments");
  BCHK(hipSetDevice(b->device));
  const int n = (int)b->in.size();
  if (use_graph && !b->gexec) {
    BCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
    int rc = run_iteration(b, true);
    hipGraph_t gr = nullptr;
    const hipError_t ec = hipStreamEndCapture(b->st, &gr);
    if (rc) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    BCHK(ec);
    b->graph = gr;
    BCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
  }
  int active = 0;
  for (int it = 0; it < niter; ++it) {
    const auto t0 = std::chrono::steady_clock::now();
    if (use_graph) {
      BCHK(hipGraphLaunch(b->gexec, b->st));
    } else {
      int rc = run_iteration(b, true);
      if (rc) return rc;
    }
    BCHK(hipMemcpyAsync(b->hskip, b->dskip, sizeof(int) * n, hipMemcpyDeviceToHost, b->st));
    BCHK(hipStreamSynchronize(b->st));
    const auto t1 = std::chrono::steady_clock::now();
    b->elapsed_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
    active = 0;
    for (int i = 0; i < n; ++i) {
      if (b->hskip[i] && b->done_ms[i] < 0) b->done_ms[i] = b->elapsed_ms;
      if (!b->hskip[i]) ++active;
    }
    if (active == 0) break;
    if (b->mo.compact && 2 * active <= b->included) {
      // at most half of the launch sequence's instances still advance: rebuild it over those
      for (int i = 0; i < n; ++i) b->use[i] = b->hskip[i] ? 0 : 1;
      if (int rc = rebuild_ops(b)) return rc;
      if (use_graph) {
        BCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
        int rc2 = run_iteration(b, true);
        hipGraph_t gr = nullptr;
        const hipError_t ec = hipStreamEndCapture(b->st, &gr);
        if (rc2) {
          if (gr) (void)hipGraphDestroy(gr);
          return rc2;
        }
        BCHK(ec);
        b->graph = gr;
        BCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
      }
    }
  }
  b->swept = true;
  if (n_active) *n_active = active;
  return WLC_OK;
}

int wlc_batch_status(wlc_batch* b, int i, wlc_scft_status* out) {
  int rc = check_inst(b, i);
  if (rc) return rc;
  if (!out) return host_fail(WLC_EINVAL, "NULL argument");
  BCHK(hipSetDevice(b->device));
  BCHK(hipStreamSynchronize(b->st));
  MixState s;
  BCHK(hipMemcpy(&s, b->dstate + i, sizeof(MixState), hipMemcpyDeviceToHost));
  std::memset(out, 0, sizeof(*out));
  out->iterations = s.iter;
  out->converged = s.converged;
  out->done = s.done;
  out->F = s.F;
  out->lnQ = s.lnQ;
  for (int a = 0; a < 3; ++a) out->resid[a] = s.resid[a];
  out->restarts = s.restarts;
  out->time_ms = b->done_ms[i];
  out->elapsed_ms = b->elapsed_ms;
  return WLC_OK;
}

}  // extern "C"
