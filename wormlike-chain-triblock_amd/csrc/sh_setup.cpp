// Host-side orientation algebra for the k-aligned MDE solve.
//
// The reference builds, per k-point, the dense complex-symmetric matrix
//   G(k) = c I + ds lam diag(l(l+1)) + i sigma ds (k_x R_x + k_y R_y + k_z R_z)
// (Gmatrix_2dcomp_mod.f90:470-495) from the real-spherical-harmonic coupling
// matrices R_a = <Y_i|u_a|Y_j> (init_array_2dcomp_mod.f90:166-182, triple product of
// SPH_mod.f90:11-98) and inverts it with LAPACK.  Here the same R_a are derived from
// the complex Gaunt coefficient and a unitary real<->complex map, and the per-k
// solve is done in the frame where k is the polar axis: there k.R = |k| R_z is
// tridiagonal in l for each m.  The rotation into that frame is written as
//   E(k) = Yh^T Zr(theta) Yh Zr(phi + pi/2)
// with Yh the (constant) representation of a pi/2 rotation about y and Zr(a) the
// representation of a rotation about z, so that G^-1 = E^T T^-1 E.
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "wlc_internal.h"

namespace wlc {
namespace {

using cplx = std::complex<double>;

double factorial(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}

// Wigner 3j symbol, Racah's single-sum formula.
double wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
  if (m1 + m2 + m3 != 0) return 0.0;
  if (j3 < std::abs(j1 - j2) || j3 > j1 + j2) return 0.0;
  if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.0;
  const double tri = factorial(j1 + j2 - j3) * factorial(j1 - j2 + j3) *
                     factorial(-j1 + j2 + j3) / factorial(j1 + j2 + j3 + 1);
  const double mm = factorial(j1 + m1) * factorial(j1 - m1) * factorial(j2 + m2) *
                    factorial(j2 - m2) * factorial(j3 + m3) * factorial(j3 - m3);
  double sum = 0.0;
  for (int t = 0; t <= j1 + j2 + j3; ++t) {
    const int a = j3 - j2 + t + m1, b = j3 - j1 + t - m2, c = j1 + j2 - j3 - t;
    const int d = j1 - t - m1, e = j2 - t + m2;
    if (a < 0 || b < 0 || c < 0 || d < 0 || e < 0) continue;
    const double den = factorial(t) * factorial(a) * factorial(b) * factorial(c) *
                       factorial(d) * factorial(e);
    sum += ((t & 1) ? -1.0 : 1.0) / den;
  }
  const int ph = j1 - j2 - m3;
  return ((ph & 1) ? -1.0 : 1.0) * std::sqrt(tri * mm) * sum;
}

// Row m of the unitary map from complex Y_{l,mu} to the real basis:
//   m > 0: (Y_{l,-m} + (-1)^m Y_{l,m}) / sqrt2
//   m < 0: i (Y_{l,m} - (-1)^m Y_{l,-m}) / sqrt2      (m -> |m| in the phase)
//   m = 0: Y_{l,0}
cplx umap(int m, int mu) {
  const double r2 = 1.0 / std::sqrt(2.0);
  if (m == 0) return mu == 0 ? cplx(1.0, 0.0) : cplx(0.0, 0.0);
  const int a = std::abs(m);
  const double sg = (a & 1) ? -1.0 : 1.0;
  if (m > 0) {
    if (mu == -a) return cplx(r2, 0.0);
    if (mu == a) return cplx(sg * r2, 0.0);
    return 0.0;
  }
  if (mu == -a) return cplx(0.0, r2);
  if (mu == a) return cplx(0.0, -sg * r2);
  return 0.0;
}

// sqrt(4 pi) * integral of three real harmonics, via the complex Gaunt coefficient
double real_gaunt(int l1, int l2, int l3, int m1, int m2, int m3) {
  const double pre = std::sqrt(double((2 * l1 + 1) * (2 * l2 + 1) * (2 * l3 + 1))) *
                     wigner3j(l1, l2, l3, 0, 0, 0);
  if (pre == 0.0) return 0.0;
  cplx s = 0.0;
  for (int a = -l1; a <= l1; ++a) {
    const cplx u1 = umap(m1, a);
    if (u1 == cplx(0.0)) continue;
    for (int b = -l2; b <= l2; ++b) {
      const cplx u2 = umap(m2, b);
      if (u2 == cplx(0.0)) continue;
      const int c = -a - b;
      if (std::abs(c) > l3) continue;
      const cplx u3 = umap(m3, c);
      if (u3 == cplx(0.0)) continue;
      s += u1 * u2 * u3 * wigner3j(l1, l2, l3, a, b, c);
    }
  }
  return pre * s.real();
}

using Mat = std::vector<double>;  // row-major M x M

// Representation D(Q) of a rotation Q on the truncated real-SH space, from the
// covariance D R_a D^T = sum_b Q_ab R_b, solved block by block in l: with
// A_l = [R_x; R_y; R_z] restricted to (l, l+1), A_l^T A_l = c_l I, so
// D_{l+1}^T = (1/c_l) sum_a R_a^{(l+1,l)} D_l^T B_a.
std::string rep_of_rotation(int L, const Mat* R, const double Q[3][3], Mat* D) {
  const int M = (L + 1) * (L + 1);
  D->assign(size_t(M) * M, 0.0);
  (*D)[0] = 1.0;
  for (int l = 0; l < L; ++l) {
    const int a0 = l * l, n0 = 2 * l + 1, b0 = (l + 1) * (l + 1), n1 = 2 * l + 3;
    // c_l from the diagonal of A^T A
    double cl = 0.0;
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < n0; ++i) {
        double v = R[a][size_t(a0 + i) * M + b0];
        cl += v * v;
      }
    if (cl <= 0.0) return "degenerate coupling block";
    std::vector<double> acc(size_t(n1) * n1, 0.0);
    std::vector<double> Ba(size_t(n0) * n1), DtB(size_t(n0) * n1);
    for (int a = 0; a < 3; ++a) {
      for (int i = 0; i < n0; ++i)
        for (int j = 0; j < n1; ++j) {
          double s = 0.0;
          for (int b = 0; b < 3; ++b) s += Q[a][b] * R[b][size_t(a0 + i) * M + b0 + j];
          Ba[size_t(i) * n1 + j] = s;
        }
      for (int i = 0; i < n0; ++i)       // DtB = D_l^T Ba
        for (int j = 0; j < n1; ++j) {
          double s = 0.0;
          for (int p = 0; p < n0; ++p) s += (*D)[size_t(a0 + p) * M + a0 + i] * Ba[size_t(p) * n1 + j];
          DtB[size_t(i) * n1 + j] = s;
        }
      for (int i = 0; i < n1; ++i)       // acc += R_a^{(l+1,l)} DtB
        for (int j = 0; j < n1; ++j) {
          double s = 0.0;
          for (int p = 0; p < n0; ++p) s += R[a][size_t(a0 + p) * M + b0 + i] * DtB[size_t(p) * n1 + j];
          acc[size_t(i) * n1 + j] += s;
        }
    }
    for (int i = 0; i < n1; ++i)
      for (int j = 0; j < n1; ++j) (*D)[size_t(b0 + i) * M + b0 + j] = acc[size_t(j) * n1 + i] / cl;
  }
  return "";
}

void rot_z(double a, double Q[3][3]) {
  const double c = std::cos(a), s = std::sin(a);
  const double q[3][3] = {{c, -s, 0}, {s, c, 0}, {0, 0, 1}};
  std::memcpy(Q, q, sizeof(q));
}
void rot_y(double a, double Q[3][3]) {
  const double c = std::cos(a), s = std::sin(a);
  const double q[3][3] = {{c, 0, s}, {0, 1, 0}, {-s, 0, c}};
  std::memcpy(Q, q, sizeof(q));
}

// apply Zr(alpha) (cos/sin of |m| alpha) in place
void apply_zr(const OrientTables& t, int L, double ca, double sa, cplx* v) {
  double cm[kMaxL + 1], sm[kMaxL + 1];
  cm[0] = 1.0; sm[0] = 0.0;
  for (int m = 1; m <= L; ++m) {
    cm[m] = cm[m - 1] * ca - sm[m - 1] * sa;
    sm[m] = sm[m - 1] * ca + cm[m - 1] * sa;
  }
  for (int l = 1; l <= L; ++l)
    for (int m = 1; m <= l; ++m) {
      const int ip = sh_index(l, m), in = sh_index(l, -m);
      const cplx vp = v[ip], vn = v[in];
      v[ip] = cm[m] * vp + t.zsign[ip] * sm[m] * vn;
      v[in] = cm[m] * vn + t.zsign[in] * sm[m] * vp;
    }
}

void apply_yh(const OrientTables& t, int L, bool transpose, cplx* v) {
  for (int l = 1; l <= L; ++l) {
    cplx in[2 * kMaxL + 1], out[2 * kMaxL + 1];
    for (int m = -l; m <= l; ++m) { in[m + l] = v[sh_index(l, m)]; out[m + l] = 0.0; }
    int e = t.yh_off[l];
    for (int m = -l; m <= l; ++m)
      for (int mp = -l; mp <= l; ++mp) {
        if (!yh_pattern(l, m, mp)) continue;
        const double c = t.yh[e++];
        if (!transpose) out[m + l] += c * in[mp + l];
        else out[mp + l] += c * in[m + l];
      }
    for (int m = -l; m <= l; ++m) v[sh_index(l, m)] = out[m + l];
  }
}

}  // namespace

void host_rotated_solve(const OrientTables& t, int L, double cdelta, double sigma, double ds,
                        double dslam, const double k[3], double* vd) {
  cplx* v = reinterpret_cast<cplx*>(vd);
  const double kap = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
  if (kap == 0.0) {
    for (int l = 0; l <= L; ++l)
      for (int m = -l; m <= l; ++m) v[sh_index(l, m)] /= cdelta + dslam * l * (l + 1);
    return;
  }
  const double kp = std::sqrt(k[0] * k[0] + k[1] * k[1]);
  const double ct = k[2] / kap, st = kp / kap;
  double cp = 1.0, sp = 0.0;
  if (kp > 0.0) { cp = k[0] / kp; sp = k[1] / kp; }
  // phi + pi/2
  const double cq = -sp, sq = cp;
  apply_zr(t, L, cq, sq, v);
  apply_yh(t, L, false, v);
  apply_zr(t, L, ct, st, v);
  apply_yh(t, L, true, v);
  // tridiagonal solve per m in the rotated frame
  const double g = sigma * ds * kap;
  for (int m = -L; m <= L; ++m) {
    const int a = std::abs(m);
    double beta[kMaxL + 1];
    cplx gg[kMaxL + 1];
    for (int l = a; l <= L; ++l) {
      const double al = cdelta + dslam * l * (l + 1);
      const cplx vl = v[sh_index(l, m)];
      if (l == a) { beta[l] = al; gg[l] = vl; continue; }
      const double bb = g * t.tri_e[sh_index(l - 1, m)];   // b_{l-1} = i bb
      beta[l] = al + bb * bb / beta[l - 1];
      const double mu = bb / beta[l - 1];                    // mu = i mu
      gg[l] = vl - cplx(0.0, mu) * gg[l - 1];
    }
    cplx y = gg[L] / beta[L];
    v[sh_index(L, m)] = y;
    for (int l = L - 1; l >= a; --l) {
      const double bb = g * t.tri_e[sh_index(l, m)];
      y = (gg[l] - cplx(0.0, bb) * y) / beta[l];
      v[sh_index(l, m)] = y;
    }
  }
  apply_yh(t, L, false, v);
  apply_zr(t, L, ct, -st, v);
  apply_yh(t, L, true, v);
  apply_zr(t, L, cq, -sq, v);
}

std::string build_orientation_tables(int Lmax, OrientTables* t, std::vector<double>* oRx,
                                     std::vector<double>* oRy, std::vector<double>* oRz) {
  const int L = kMaxL;  // tables are built once for the largest supported L
  const int M = (L + 1) * (L + 1);
  std::memset(t, 0, sizeof(*t));
  Mat R[3];
  for (int a = 0; a < 3; ++a) R[a].assign(size_t(M) * M, 0.0);
  // R_x = -T(.,.,1;.,.,1)/sqrt3, R_y = -T(.,.,1;.,.,-1)/sqrt3, R_z = T(.,.,1;.,.,0)/sqrt3
  const double s3 = 1.0 / std::sqrt(3.0);
  for (int li = 0; li <= L; ++li)
    for (int mi = -li; mi <= li; ++mi)
      for (int lj = std::max(0, li - 1); lj <= std::min(L, li + 1); ++lj)
        for (int mj = -lj; mj <= lj; ++mj) {
          const int i = sh_index(li, mi), j = sh_index(lj, mj);
          R[0][size_t(i) * M + j] = -s3 * real_gaunt(li, lj, 1, mi, mj, 1);
          R[1][size_t(i) * M + j] = -s3 * real_gaunt(li, lj, 1, mi, mj, -1);
          R[2][size_t(i) * M + j] = s3 * real_gaunt(li, lj, 1, mi, mj, 0);
        }
  // z rotation structure
  double Q[3][3];
  const double alpha = 0.7;
  rot_z(alpha, Q);
  Mat Z;
  std::string err = rep_of_rotation(L, R, Q, &Z);
  if (!err.empty()) return err;
  for (int l = 0; l <= L; ++l)
    for (int m = -l; m <= l; ++m) {
      const int i = sh_index(l, m);
      for (int j = 0; j < M; ++j) {
        const double zij = Z[size_t(i) * M + j];
        if (j == i) {
          if (std::fabs(zij - std::cos(std::abs(m) * alpha)) > 1e-11) return "unexpected z-rotation diagonal";
        } else if (m != 0 && j == sh_index(l, -m)) {
          const double sgn = zij / std::sin(std::abs(m) * alpha);
          if (std::fabs(std::fabs(sgn) - 1.0) > 1e-11) return "unexpected z-rotation pairing";
          t->zsign[i] = sgn > 0 ? 1.0 : -1.0;
        } else if (std::fabs(zij) > 1e-11) {
          return "z-rotation couples outside (m, -m) pairs";
        }
      }
    }
  // pi/2 rotation about y, packed on its structural pattern
  rot_y(M_PI / 2, Q);
  Mat Y;
  err = rep_of_rotation(L, R, Q, &Y);
  if (!err.empty()) return err;
  int e = 0;
  for (int l = 0; l <= L; ++l) {
    t->yh_off[l] = e;
    for (int m = -l; m <= l; ++m)
      for (int mp = -l; mp <= l; ++mp) {
        const double y = Y[size_t(sh_index(l, m)) * M + sh_index(l, mp)];
        if (yh_pattern(l, m, mp)) t->yh[e++] = y;
        else if (std::fabs(y) > 1e-11) return "pi/2 y-rotation outside its parity pattern";
      }
  }
  t->yh_off[L + 1] = e;
  // rotated-frame tridiagonal couplings; R_z must preserve m
  for (int l = 0; l < L; ++l)
    for (int m = -l; m <= l; ++m) {
      t->tri_e[sh_index(l, m)] = R[2][size_t(sh_index(l, m)) * M + sh_index(l + 1, m)];
      if (std::fabs(t->tri_e[sh_index(l, m)] - R[2][size_t(sh_index(l, -m)) * M + sh_index(l + 1, -m)]) > 1e-14)
        return "R_z coupling differs between m and -m";
    }
  // self-check: rotated solve == dense complex solve of G at a few k
  const int Lc = std::min(std::max(Lmax, 1), L);
  const int Mc = (Lc + 1) * (Lc + 1);
  const double kk[4][3] = {{3.1, -1.7, 2.2}, {0.0, 0.0, 4.0}, {-2.5, 0.0, 0.0}, {1.0, 7.0, -0.3}};
  for (int trial = 0; trial < 4; ++trial) {
    const double cdelta = 11.0 / 6.0, sigma = trial & 1 ? -1.0 : 1.0, ds = 0.01, dslam = 0.013;
    std::vector<cplx> G(size_t(Mc) * Mc), b(Mc), x(Mc);
    for (int i = 0; i < Mc; ++i) {
      int l = (int)std::floor(std::sqrt((double)i));
      b[i] = cplx(std::sin(1.0 + i), std::cos(2.0 * i));
      for (int j = 0; j < Mc; ++j) {
        double im = 0.0;
        for (int a = 0; a < 3; ++a) im += kk[trial][a] * R[a][size_t(i) * M + j];
        G[size_t(i) * Mc + j] = cplx(i == j ? cdelta + dslam * l * (l + 1) : 0.0, sigma * ds * im);
      }
    }
    // Gaussian elimination with partial pivoting
    std::vector<cplx> A = G, rhs = b;
    for (int c = 0; c < Mc; ++c) {
      int p = c;
      for (int r = c + 1; r < Mc; ++r)
        if (std::abs(A[size_t(r) * Mc + c]) > std::abs(A[size_t(p) * Mc + c])) p = r;
      if (p != c) {
        for (int j = 0; j < Mc; ++j) std::swap(A[size_t(c) * Mc + j], A[size_t(p) * Mc + j]);
        std::swap(rhs[c], rhs[p]);
      }
      for (int r = c + 1; r < Mc; ++r) {
        const cplx f = A[size_t(r) * Mc + c] / A[size_t(c) * Mc + c];
        for (int j = c; j < Mc; ++j) A[size_t(r) * Mc + j] -= f * A[size_t(c) * Mc + j];
        rhs[r] -= f * rhs[c];
      }
    }
    for (int r = Mc - 1; r >= 0; --r) {
      cplx s = rhs[r];
      for (int j = r + 1; j < Mc; ++j) s -= A[size_t(r) * Mc + j] * x[j];
      x[r] = s / A[size_t(r) * Mc + r];
    }
    std::vector<cplx> v = b;
    host_rotated_solve(*t, Lc, cdelta, sigma, ds, dslam, kk[trial], reinterpret_cast<double*>(v.data()));
    double num = 0.0, den = 0.0;
    for (int i = 0; i < Mc; ++i) { num = std::max(num, std::abs(v[i] - x[i])); den = std::max(den, std::abs(x[i])); }
    if (!(num <= 1e-12 * den)) return "rotated-frame solve self-check failed";
  }
  if (oRx) *oRx = R[0];
  if (oRy) *oRy = R[1];
  if (oRz) *oRz = R[2];
  return "";
}

std::string make_axis_plan(int n, AxisPlan* p) {
  if (n < 2) return "grid edge must be >= 2";
  p->n = n;
  p->npass = 0;
  int r = n, a2 = 0;
  while (r % 2 == 0) { r /= 2; ++a2; }
  int a3 = 0;
  while (r % 3 == 0) { r /= 3; ++a3; }
  if (r != 1) return "grid edges must factor into 2s and 3s";
  // powers of two: radix 8 where possible, a 4 or two 4s to avoid radix-2 passes
  int n8 = a2 / 3, rem = a2 % 3;
  if (rem == 1 && n8 > 0) { n8 -= 1; rem = 4; }
  for (int i = 0; i < n8; ++i) p->radix[p->npass++] = 8;
  if (rem == 4) { p->radix[p->npass++] = 4; p->radix[p->npass++] = 4; }
  else if (rem == 2) p->radix[p->npass++] = 4;
  else if (rem == 1) p->radix[p->npass++] = 2;
  for (int i = 0; i < a3; ++i) p->radix[p->npass++] = 3;
  p->tw.resize(2 * size_t(n));
  const long double tp = 2.0L * acosl(-1.0L);
  for (int t = 0; t < n; ++t) {
    const long double ang = -tp * (long double)t / (long double)n;
    p->tw[2 * t] = (double)cosl(ang);
    p->tw[2 * t + 1] = (double)sinl(ang);
  }
  return "";
}

}  // namespace wlc
