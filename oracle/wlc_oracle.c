/*
 * wlc_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the wormlike-chain MDE
 * propagator solve of tjz2026/wormlike-chain-triblock.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
 * code, and only as the checker or the timed CPU baseline; the product library
 * never links or calls it.  It restates the reference's algorithm in plain C:
 *
 *   - Wigner 3j and the real-SH triple product ............ SPH_mod.f90:11-142
 *   - basis ordering i <-> (l, m), l=0..L, m=-l..l ........ SPH_mod.f90:192-222
 *   - R_x, R_y, R_z = -/+ T(l_i,l_j,1;m_i,m_j,+-1|0)/sqrt3 . init_array_2dcomp_mod.f90:166-182
 *   - G = c I + ds lam l(l+1) + i sigma ds k.R, dense inverse, lower triangle
 *     kept and applied as a symmetric matrix ............ Gmatrix_2dcomp_mod.f90:384-543,
 *                                                          matrix_inverse_mod.f90:10-33,85-124,
 *                                                          wormlike_MDE_2dcomp_mod.f90:395-409
 *   - signed k mapping, Nyquist counted +n/2 ............. Gmatrix_2dcomp_mod.f90:450-467
 *   - BDF1/BDF2/BDF3 right-hand sides and junctions ...... wormlike_MDE_2dcomp_mod.f90:323-937 (forward),
 *                                                          :1035-1532 (backward)
 *   - R2C forward / C2R inverse FFT, x axis halved, /N ... wormlike_MDE_2dcomp_mod.f90:357-455
 *   - Q, per-block contour Simpson, totden, phi .......... SCFT_Density_FE.f90:48-159,
 *                                                          SCFT_Density_FE_ABC.f90:98-139
 *   - composite Simpson / periodic 3D Simpson ............ utility_mod.f90:238-298,346-374
 *   - free energy (AB form generalised to ABC) ........... SCFT_Density_FE.f90:273-333
 *
 * Two arithmetic modes: literal=0 is intended FP64; literal=1 additionally rounds
 * the reference's default-real literals (11/6, 1/3, 1/sqrt3, pi, the sqrt inside the
 * triple product) and the G-applied spectrum to FP32, as the Fortran does.
 *
 * Build decisions the reference leaves open (documented in the repository's design
 * notes): three blocks; junction target point takes the lower block's field in
 * both directions; BDF startup only at the chain ends; odd block lengths use
 * Simpson plus a 3/8 tail; spectra follow numpy's irfftn semantics at the
 * self-conjugate x planes.
 */
#include "wlc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define MAXL 16

/* ------------------------------------------------------------------------- */
/* SH algebra (SPH_mod.f90)                                                   */
/* ------------------------------------------------------------------------- */

static int sign_ov(int i) { return (i % 2 == 0) ? 1 : -1; } /* SPH_mod.f90:146-156 */

static double fact(int l) { /* SPH_mod.f90:158-171 */
  double f = 1.0;
  for (int i = 2; i <= l; i++) f *= i;
  return f;
}

static double fact_con(int m1, int m2) { /* SPH_mod.f90:174-189 */
  double f = 1.0;
  for (int i = m1; i <= m2; i++) f *= i;
  return f;
}

static int imax3(int a, int b, int c) { int m = a > b ? a : b; return m > c ? m : c; }
static int imin3(int a, int b, int c) { int m = a < b ? a : b; return m < c ? m : c; }

/* Racah formula, SPH_mod.f90:101-142 (same operation order) */
double orc_three_j(int l1, int l2, int l3, int m1, int m2, int m3) {
  const int FACTVALUE = 20;
  if ((m1 + m2 + m3) != 0 || (l1 + l2 - l3) < 0 || (l1 - l2 + l3) < 0 || (l2 + l3 - l1) < 0 ||
      abs(l1) < abs(m1) || abs(l2) < abs(m2) || abs(l3) < abs(m3))
    return 0.0;
  double temp1, temp2;
  if ((l1 + l2 + l3 + 1) > FACTVALUE) {
    temp1 = 1.0 / fact(FACTVALUE);
    temp2 = 1.0 / fact_con(FACTVALUE + 1, l1 + l2 + l3 + 1);
  } else {
    temp1 = 1.0 / fact(l1 + l2 + l3 + 1);
    temp2 = 1.0;
  }
  double temp = fact(l1 + l2 - l3) * fact(l1 - l2 + l3) * fact(l2 + l3 - l1) * temp1;
  temp = temp * fact(l1 - m1) * fact(l1 + m1) * fact(l2 - m2) * fact(m2 + l2) * fact(l3 - m3) *
         fact(l3 + m3) * temp2;
  temp = sign_ov(l1 - l2 - m3) * sqrt(temp);
  int kmin = imax3(0, l2 - l3 - m1, l1 - l3 + m2);
  int kmax = imin3(l1 + l2 - l3, l1 - m1, l2 + m2);
  double suma = 0.0;
  for (int k = kmin; k <= kmax; k++) {
    double tk = temp / fact(k);
    tk = tk / fact(l1 + l2 - l3 - k);
    tk = tk / fact(l1 - m1 - k);
    tk = tk / fact(l2 + m2 - k);
    tk = tk / fact(l3 - l2 + m1 + k);
    tk = tk / fact(l3 - l1 - m2 + k);
    suma = suma + sign_ov(k) * tk;
  }
  return suma;
}

/* sqrt of the integer product (2l1+1)(2l2+1)(2l3+1) [/2]: the reference takes it in
 * default (single) precision, SPH_mod.f90:25-26,33,37,44,... */
static double sqrt_prod(int l1, int l2, int l3, int half, int literal) {
  if (literal) {
    float p = 1.0f * (float)(2 * l1 + 1) * (float)(2 * l2 + 1) * (float)(2 * l3 + 1);
    if (half) p = p / 2.0f;
    return (double)sqrtf(p);
  }
  double p = (double)(2 * l1 + 1) * (double)(2 * l2 + 1) * (double)(2 * l3 + 1);
  if (half) p = p / 2.0;
  return sqrt(p);
}

/* real-SH triple product x sqrt(4 pi), SPH_mod.f90:11-98 (cases 1-11) */
double orc_triple_product(int l1, int l2, int l3, int m1, int m2, int m3, int literal) {
  double sqrt1 = sqrt_prod(l1, l2, l3, 0, literal);
  double sqrt2 = literal ? sqrt1 / (double)sqrtf(2.0f) : sqrt1 / sqrt(2.0);
  double trpd = 0.0, temp, suma;
  if (sign_ov(l1 + l2 + l3) == -1) return 0.0;
  if (m1 == 0 && m2 == 0 && m3 == 0) { /* 1 */
    temp = orc_three_j(l1, l2, l3, 0, 0, 0);
    trpd = sqrt_prod(l1, l2, l3, 0, literal) * temp * temp;
  } else if (m1 > 0 && m2 > 0 && m3 > 0) { /* 2 */
    temp = sqrt_prod(l1, l2, l3, 1, literal);
    temp = temp * orc_three_j(l1, l2, l3, 0, 0, 0);
    suma = sign_ov(m1) * orc_three_j(l1, l2, l3, -m1, m2, m3) +
           sign_ov(m2) * orc_three_j(l1, l2, l3, m1, -m2, m3) +
           sign_ov(m3) * orc_three_j(l1, l2, l3, m1, m2, -m3);
    trpd = temp * suma;
  } else if (m1 == 0 && m2 > 0 && m3 > 0) { /* 3 */
    temp = sign_ov(m2) * sqrt_prod(l1, l2, l3, 0, literal);
    trpd = temp * orc_three_j(l1, l2, l3, 0, 0, 0) * orc_three_j(l1, l2, l3, 0, -m2, m3);
  } else if (m1 > 0 && m2 == 0 && m3 > 0) { /* 4 */
    temp = sign_ov(m1) * sqrt_prod(l1, l2, l3, 0, literal);
    trpd = temp * orc_three_j(l1, l2, l3, 0, 0, 0) * orc_three_j(l1, l2, l3, -m1, 0, m3);
  } else if (m1 > 0 && m2 > 0 && m3 == 0) { /* 5 */
    temp = sign_ov(m1) * sqrt_prod(l1, l2, l3, 0, literal);
    trpd = temp * orc_three_j(l1, l2, l3, 0, 0, 0) * orc_three_j(l1, l2, l3, -m1, m2, 0);
  } else if (m1 == 0 && m2 < 0 && m3 < 0) { /* 6 */
    temp = sign_ov(m2) * sqrt1;
    trpd = temp * orc_three_j(l1, l2, l3, 0, 0, 0) * orc_three_j(l1, l2, l3, 0, m2, -m3);
  } else if (m1 < 0 && m2 == 0 && m3 < 0) { /* 7 */
    temp = sign_ov(m1) * sqrt1;
    trpd = temp * orc_three_j(l1, l2, l3, 0, 0, 0) * orc_three_j(l1, l2, l3, m1, 0, -m3);
  } else if (m1 < 0 && m2 < 0 && m3 == 0) { /* 8 */
    temp = sign_ov(m1) * sqrt1;
    trpd = temp * orc_three_j(l1, l2, l3, 0, 0, 0) * orc_three_j(l1, l2, l3, m1, -m2, 0);
  } else if (m1 > 0 && m2 < 0 && m3 < 0) { /* 9 */
    temp = sqrt2 * orc_three_j(l1, l2, l3, 0, 0, 0);
    suma = sign_ov(m2) * orc_three_j(l1, l2, l3, m1, m2, -m3) +
           sign_ov(m3) * orc_three_j(l1, l2, l3, m1, -m2, m3) -
           sign_ov(m1) * orc_three_j(l1, l2, l3, m1, m2, m3);
    trpd = temp * suma;
  } else if (m1 < 0 && m2 > 0 && m3 < 0) { /* 10 */
    temp = sqrt2 * orc_three_j(l1, l2, l3, 0, 0, 0);
    suma = sign_ov(m1) * orc_three_j(l1, l2, l3, m1, m2, -m3) -
           sign_ov(m2) * orc_three_j(l1, l2, l3, m1, m2, m3) +
           sign_ov(m3) * orc_three_j(l1, l2, l3, -m1, m2, m3);
    trpd = temp * suma;
  } else if (m1 < 0 && m2 < 0 && m3 > 0) { /* 11 */
    temp = sqrt2 * orc_three_j(l1, l2, l3, 0, 0, 0);
    suma = sign_ov(m1) * orc_three_j(l1, l2, l3, m1, -m2, m3) +
           sign_ov(m2) * orc_three_j(l1, l2, l3, -m1, m2, m3) -
           sign_ov(m3) * orc_three_j(l1, l2, l3, m1, m2, m3);
    trpd = temp * suma;
  } else {
    trpd = 0.0;
  }
  return trpd;
}

static void basis_lm(int i, int* l, int* m) { /* SPH_mod.f90:214-222: i = l*l + l + m */
  int ll = (int)floor(sqrt((double)i));
  while ((ll + 1) * (ll + 1) <= i) ll++;
  while (ll * ll > i) ll--;
  *l = ll;
  *m = i - ll * ll - ll;
}

/* init_array_2dcomp_mod.f90:166-182 */
void orc_rmatrices(int L, int literal, double* Rx, double* Ry, double* Rz) {
  int M = (L + 1) * (L + 1);
  double tneg = literal ? (double)(-1.0f / sqrtf(3.0f)) : -1.0 / sqrt(3.0);
  double tpos = literal ? (double)(1.0f / sqrtf(3.0f)) : 1.0 / sqrt(3.0);
  for (int i = 0; i < M; i++) {
    int li, mi; basis_lm(i, &li, &mi);
    for (int j = 0; j < M; j++) {
      int lj, mj; basis_lm(j, &lj, &mj);
      Rx[i * M + j] = tneg * orc_triple_product(li, lj, 1, mi, mj, 1, literal);
      Ry[i * M + j] = tneg * orc_triple_product(li, lj, 1, mi, mj, -1, literal);
      Rz[i * M + j] = tpos * orc_triple_product(li, lj, 1, mi, mj, 0, literal);
    }
  }
}

/* ------------------------------------------------------------------------- */
/* dense complex inverse (restates zgetrf/zgetri, matrix_inverse_mod.f90:85-124) */
/* ------------------------------------------------------------------------- */

/* Gauss-Jordan with partial pivoting; a is n*n complex interleaved, row-major */
static int zinv(int n, double* a) {
  int* perm = (int*)malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) perm[i] = i;
  /* augment implicitly: we do in-place Gauss-Jordan with column swaps recorded */
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = hypot(a[2 * (k * n + k)], a[2 * (k * n + k) + 1]);
    for (int i = k + 1; i < n; i++) {
      double v = hypot(a[2 * (i * n + k)], a[2 * (i * n + k) + 1]);
      if (v > best) { best = v; p = i; }
    }
    if (best == 0.0) { free(perm); return -1; }
    if (p != k) {
      for (int j = 0; j < n; j++) {
        double tr = a[2 * (k * n + j)], ti = a[2 * (k * n + j) + 1];
        a[2 * (k * n + j)] = a[2 * (p * n + j)];
        a[2 * (k * n + j) + 1] = a[2 * (p * n + j) + 1];
        a[2 * (p * n + j)] = tr;
        a[2 * (p * n + j) + 1] = ti;
      }
      int t = perm[k]; perm[k] = perm[p]; perm[p] = t;
    }
    double pr = a[2 * (k * n + k)], pi = a[2 * (k * n + k) + 1];
    double den = pr * pr + pi * pi;
    double ir = pr / den, ii = -pi / den; /* 1/pivot */
    a[2 * (k * n + k)] = 1.0;
    a[2 * (k * n + k) + 1] = 0.0;
    for (int j = 0; j < n; j++) {
      double xr = a[2 * (k * n + j)], xi = a[2 * (k * n + j) + 1];
      a[2 * (k * n + j)] = xr * ir - xi * ii;
      a[2 * (k * n + j) + 1] = xr * ii + xi * ir;
    }
    for (int i = 0; i < n; i++) {
      if (i == k) continue;
      double fr = a[2 * (i * n + k)], fi = a[2 * (i * n + k) + 1];
      if (fr == 0.0 && fi == 0.0) continue;
      a[2 * (i * n + k)] = 0.0;
      a[2 * (i * n + k) + 1] = 0.0;
      for (int j = 0; j < n; j++) {
        double xr = a[2 * (k * n + j)], xi = a[2 * (k * n + j) + 1];
        a[2 * (i * n + j)] -= fr * xr - fi * xi;
        a[2 * (i * n + j) + 1] -= fr * xi + fi * xr;
      }
    }
  }
  /* undo the row permutation of the input: columns of the inverse are permuted */
  double* tmp = (double*)malloc(sizeof(double) * 2 * n * n);
  memcpy(tmp, a, sizeof(double) * 2 * n * n);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < n; k++) {
      a[2 * (i * n + perm[k])] = tmp[2 * (i * n + k)];
      a[2 * (i * n + perm[k]) + 1] = tmp[2 * (i * n + k) + 1];
    }
  free(tmp);
  free(perm);
  return 0;
}

static double pi_value(int literal) {
  if (literal) return (double)3.14159265358979323846f; /* nr_files.f90:10 (REAL(SP)) */
  return 3.14159265358979323846264338327950288;
}

static double kfactor(int n_edge, double box, int literal) {
  /* Gmatrix_2dcomp_mod.f90:378-380: 2*PI/(SIDE*d), PI single precision */
  (void)n_edge;
  if (literal) return (double)(2.0f * 3.14159265358979323846f) / box;
  return 2.0 * pi_value(0) / box;
}

static int signed_k(int j, int n) { return (j <= n / 2) ? j : j - n; } /* :450-467 */

int orc_ginv(int L, double cdelta, double sigma, double ds, double lam, double kx, double ky,
             double kz, int literal, double* out) {
  int M = (L + 1) * (L + 1);
  double* Rx = (double*)malloc(sizeof(double) * M * M * 3);
  double* Ry = Rx + M * M;
  double* Rz = Ry + M * M;
  orc_rmatrices(L, literal, Rx, Ry, Rz);
  for (int i = 0; i < M; i++) {
    int l, m; basis_lm(i, &l, &m);
    for (int j = 0; j < M; j++) {
      out[2 * (i * M + j)] = (i == j) ? cdelta + ds * lam * l * (l + 1) : 0.0;
      out[2 * (i * M + j) + 1] =
          sigma * ds * (Rx[i * M + j] * kx + Ry[i * M + j] * ky + Rz[i * M + j] * kz);
    }
  }
  free(Rx);
  return zinv(M, out);
}

/* ------------------------------------------------------------------------- */
/* FFT: Stockham mixed radix, generic small DFTs                               */
/* ------------------------------------------------------------------------- */

typedef struct {
  int n, nf, f[40];
  double* tw;  /* exp(-2 pi i t / n), t < n */
} fplan;

static void fplan_init(fplan* p, int n) {
  p->n = n;
  p->nf = 0;
  int r = n;
  while (r % 4 == 0) { p->f[p->nf++] = 4; r /= 4; }
  while (r % 2 == 0) { p->f[p->nf++] = 2; r /= 2; }
  for (int q = 3; r > 1; q += 2)
    while (r % q == 0) { p->f[p->nf++] = q; r /= q; }
  p->tw = (double*)malloc(sizeof(double) * 2 * (n > 0 ? n : 1));
  long double tp = 2.0L * acosl(-1.0L);
  for (int t = 0; t < n; t++) {
    long double a = -tp * (long double)t / (long double)n;
    p->tw[2 * t] = (double)cosl(a);
    p->tw[2 * t + 1] = (double)sinl(a);
  }
}

static void fplan_free(fplan* p) { free(p->tw); }

/* x (2n) in/out, work (2n) scratch; sign -1 forward, +1 inverse */
static void fexec(const fplan* p, double* x, double* work, int sign) {
  int n = p->n;
  if (n <= 1) return;
  double* src = x;
  double* dst = work;
  int Ns = 1;
  for (int pass = 0; pass < p->nf; pass++) {
    int R = p->f[pass];
    int nb = n / R;
    double vr[64], vi[64], orr[64], oi[64];
    for (int j = 0; j < nb; j++) {
      int k = j % Ns;
      for (int r = 0; r < R; r++) {
        double a = src[2 * (j + r * nb)], b = src[2 * (j + r * nb) + 1];
        int t = (int)(((long long)k * r * (n / (Ns * R))) % n);
        /* tw[t] = exp(-2 pi i t/n); the inverse uses its conjugate */
        double c = p->tw[2 * t];
        double s = (sign < 0) ? p->tw[2 * t + 1] : -p->tw[2 * t + 1];
        vr[r] = a * c - b * s;
        vi[r] = a * s + b * c;
      }
      /* DFT_R: out[u] = sum_r v[r] exp(sign 2 pi i u r / R) */
      for (int u = 0; u < R; u++) {
        double sr = 0.0, si = 0.0;
        for (int r = 0; r < R; r++) {
          int t = (int)(((long long)u * r % R) * (n / R));
          double c = p->tw[2 * t];
          double s = (sign < 0) ? p->tw[2 * t + 1] : -p->tw[2 * t + 1];
          sr += vr[r] * c - vi[r] * s;
          si += vr[r] * s + vi[r] * c;
        }
        orr[u] = sr;
        oi[u] = si;
      }
      int idxD = (j / Ns) * Ns * R + k;
      for (int u = 0; u < R; u++) {
        dst[2 * (idxD + u * Ns)] = orr[u];
        dst[2 * (idxD + u * Ns) + 1] = oi[u];
      }
    }
    double* t = src; src = dst; dst = t;
    Ns *= R;
  }
  if (src != x) memcpy(x, src, sizeof(double) * 2 * n);
}

void orc_fft1(int n, double* x, int sign) {
  fplan p;
  fplan_init(&p, n);
  double* w = (double*)malloc(sizeof(double) * 2 * n);
  fexec(&p, x, w, sign);
  free(w);
  fplan_free(&p);
}

typedef struct {
  int nx, ny, nz, nxh;
  fplan px, py, pz;
} fft3;

static void fft3_init(fft3* f, int nx, int ny, int nz) {
  f->nx = nx; f->ny = ny; f->nz = nz; f->nxh = nx / 2 + 1;
  fplan_init(&f->px, nx);
  fplan_init(&f->py, ny);
  fplan_init(&f->pz, nz);
}

static void fft3_free(fft3* f) { fplan_free(&f->px); fplan_free(&f->py); fplan_free(&f->pz); }

/* R2C: x rows, then y, then z (2dcomp_fft_mod.f90:39-73: X pencil real -> Z pencil) */
static void fft3_r2c(const fft3* f, const double* in, double* out) {
  int nx = f->nx, ny = f->ny, nz = f->nz, nxh = f->nxh;
  int nmax = nx > ny ? nx : ny; if (nz > nmax) nmax = nz;
#pragma omp parallel
  {
    double* b = (double*)malloc(sizeof(double) * 2 * nmax);
    double* w = (double*)malloc(sizeof(double) * 2 * nmax);
#pragma omp for schedule(static)
    for (int zy = 0; zy < nz * ny; zy++) {
      const double* row = in + (size_t)zy * nx;
      for (int x = 0; x < nx; x++) { b[2 * x] = row[x]; b[2 * x + 1] = 0.0; }
      fexec(&f->px, b, w, -1);
      memcpy(out + (size_t)2 * zy * nxh, b, sizeof(double) * 2 * nxh);
    }
#pragma omp for schedule(static)
    for (int zk = 0; zk < nz * nxh; zk++) {
      int z = zk / nxh, kx = zk % nxh;
      for (int y = 0; y < ny; y++) {
        size_t o = 2 * (((size_t)z * ny + y) * nxh + kx);
        b[2 * y] = out[o]; b[2 * y + 1] = out[o + 1];
      }
      fexec(&f->py, b, w, -1);
      for (int y = 0; y < ny; y++) {
        size_t o = 2 * (((size_t)z * ny + y) * nxh + kx);
        out[o] = b[2 * y]; out[o + 1] = b[2 * y + 1];
      }
    }
#pragma omp for schedule(static)
    for (int yk = 0; yk < ny * nxh; yk++) {
      for (int z = 0; z < nz; z++) {
        size_t o = 2 * ((size_t)z * ny * nxh + yk);
        b[2 * z] = out[o]; b[2 * z + 1] = out[o + 1];
      }
      fexec(&f->pz, b, w, -1);
      for (int z = 0; z < nz; z++) {
        size_t o = 2 * ((size_t)z * ny * nxh + yk);
        out[o] = b[2 * z]; out[o + 1] = b[2 * z + 1];
      }
    }
    free(b); free(w);
  }
}

/* C2R (unnormalised): inverse z, inverse y, then x c2r; the imaginary parts of the
 * kx=0 and kx=nx/2 entries are ignored (Hermitian extension along x). spec is
 * destroyed. */
static void fft3_c2r(const fft3* f, double* spec, double* out) {
  int nx = f->nx, ny = f->ny, nz = f->nz, nxh = f->nxh;
  int nmax = nx > ny ? nx : ny; if (nz > nmax) nmax = nz;
#pragma omp parallel
  {
    double* b = (double*)malloc(sizeof(double) * 2 * nmax);
    double* w = (double*)malloc(sizeof(double) * 2 * nmax);
#pragma omp for schedule(static)
    for (int yk = 0; yk < ny * nxh; yk++) {
      for (int z = 0; z < nz; z++) {
        size_t o = 2 * ((size_t)z * ny * nxh + yk);
        b[2 * z] = spec[o]; b[2 * z + 1] = spec[o + 1];
      }
      fexec(&f->pz, b, w, +1);
      for (int z = 0; z < nz; z++) {
        size_t o = 2 * ((size_t)z * ny * nxh + yk);
        spec[o] = b[2 * z]; spec[o + 1] = b[2 * z + 1];
      }
    }
#pragma omp for schedule(static)
    for (int zk = 0; zk < nz * nxh; zk++) {
      int z = zk / nxh, kx = zk % nxh;
      for (int y = 0; y < ny; y++) {
        size_t o = 2 * (((size_t)z * ny + y) * nxh + kx);
        b[2 * y] = spec[o]; b[2 * y + 1] = spec[o + 1];
      }
      fexec(&f->py, b, w, +1);
      for (int y = 0; y < ny; y++) {
        size_t o = 2 * (((size_t)z * ny + y) * nxh + kx);
        spec[o] = b[2 * y]; spec[o + 1] = b[2 * y + 1];
      }
    }
#pragma omp for schedule(static)
    for (int zy = 0; zy < nz * ny; zy++) {
      const double* s = spec + (size_t)2 * zy * nxh;
      for (int k = 0; k < nxh; k++) { b[2 * k] = s[2 * k]; b[2 * k + 1] = s[2 * k + 1]; }
      b[1] = 0.0;
      if (nx % 2 == 0) b[2 * (nx / 2) + 1] = 0.0;
      for (int k = nxh; k < nx; k++) { b[2 * k] = b[2 * (nx - k)]; b[2 * k + 1] = -b[2 * (nx - k) + 1]; }
      fexec(&f->px, b, w, +1);
      double* row = out + (size_t)zy * nx;
      for (int x = 0; x < nx; x++) row[x] = b[2 * x];
    }
    free(b); free(w);
  }
}

void orc_rfft3(int nx, int ny, int nz, const double* in, double* out) {
  fft3 f; fft3_init(&f, nx, ny, nz);
  fft3_r2c(&f, in, out);
  fft3_free(&f);
}

void orc_irfft3(int nx, int ny, int nz, const double* in, double* out) {
  fft3 f; fft3_init(&f, nx, ny, nz);
  size_t ns = (size_t)2 * nz * ny * (nx / 2 + 1);
  double* tmp = (double*)malloc(sizeof(double) * ns);
  memcpy(tmp, in, sizeof(double) * ns);
  fft3_c2r(&f, tmp, out);
  free(tmp);
  fft3_free(&f);
}

/* ------------------------------------------------------------------------- */
/* quadrature (utility_mod.f90)                                                */
/* ------------------------------------------------------------------------- */

/* utility_mod.f90:346-374 for even n; odd n >= 3 closes with a 3/8 panel over the
 * last three intervals; n == 1 is the trapezoid. */
void orc_simpson_weights(int n, double ds, double* wts) {
  for (int j = 0; j <= n; j++) wts[j] = 0.0;
  if (n <= 0) return;
  if (n == 1) { wts[0] = wts[1] = ds / 2.0; return; }
  int ns = (n % 2 == 0) ? n : n - 3;
  if (ns > 0) {
    for (int j = 0; j <= ns; j++) {
      double c = (j == 0 || j == ns) ? 1.0 : ((j % 2) ? 4.0 : 2.0);
      wts[j] += (ds / 3.0) * c;
    }
  }
  if (n % 2) {
    double c38 = 3.0 * ds / 8.0;
    wts[ns] += c38;
    wts[ns + 1] += 3.0 * c38;
    wts[ns + 2] += 3.0 * c38;
    wts[ns + 3] += c38;
  }
}

/* the reference's composite Simpson sum (even n), with its own operation order */
static double simpson_1d_even(int n, double ds, const double* a) {
  double FUN_A = a[0], FUN_B = a[n], FUN2 = 0.0, FUN4 = a[1];
  int s1 = 1;
  for (int flag = 1; flag <= n / 2 - 1; flag++) {
    s1++; FUN2 += a[s1];
    s1++; FUN4 += a[s1];
  }
  return (ds / 3.0) * (FUN_A + FUN_B + 4.0 * FUN4 + 2.0 * FUN2);
}

static double simpson_1d(int n, double ds, const double* a) {
  if (n >= 2 && n % 2 == 0) return simpson_1d_even(n, ds, a);
  double w[4096];
  double* wp = (n + 1 <= 4096) ? w : (double*)malloc(sizeof(double) * (n + 1));
  orc_simpson_weights(n, ds, wp);
  double s = 0.0;
  for (int j = 0; j <= n; j++) s += wp[j] * a[j];
  if (wp != w) free(wp);
  return s;
}

/* utility_mod.f90:195-298 called as simposon_3D_1D_mpi(nx,ny,nz,dx,dy,dz,f):
 * periodic rule along x, then y, then z; weights 2/3 (even index), 4/3 (odd) */
double orc_volume_integral(int nx, int ny, int nz, const double* box, const double* f) {
  double dx = box[0] / nx, dy = box[1] / ny, dz = box[2] / nz;
  double* a2 = (double*)malloc(sizeof(double) * ny * nz);
  double* a1 = (double*)malloc(sizeof(double) * nz);
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++) {
      const double* a = f + ((size_t)k * ny + j) * nx;
      double FUN_A = a[0], FUN_B = a[0], FUN2 = 0.0, FUN4 = a[1];
      int s1 = 1;
      for (int flag = 1; flag <= nx / 2 - 1; flag++) {
        s1++; FUN2 += a[s1];
        s1++; FUN4 += a[s1];
      }
      a2[k * ny + j] = (dx / 3.0) * (FUN_A + FUN_B + 4.0 * FUN4 + 2.0 * FUN2);
    }
  for (int k = 0; k < nz; k++) {
    const double* a = a2 + (size_t)k * ny;
    double FUN_A = a[0], FUN_B = a[0], FUN2 = 0.0, FUN4 = a[1];
    int s1 = 1;
    for (int flag = 1; flag <= ny / 2 - 1; flag++) {
      s1++; FUN2 += a[s1];
      s1++; FUN4 += a[s1];
    }
    a1[k] = (dy / 3.0) * (FUN_A + FUN_B + 4.0 * FUN4 + 2.0 * FUN2);
  }
  double FUN2 = 0.0, FUN4 = 0.0;
  for (int s1 = 0; s1 < nz; s1 += 2) { FUN2 += a1[s1]; FUN4 += a1[s1 + 1]; }
  free(a2); free(a1);
  return (dz / 3.0) * (4.0 * FUN4 + 2.0 * FUN2);
}

/* ------------------------------------------------------------------------- */
/* context and sweeps                                                          */
/* ------------------------------------------------------------------------- */

/* one G^-1 table: Nk k-points x packed lower triangle (M(M+1)/2 complex) */
typedef struct {
  int valid, M;
  double* lt;
} gtab;

struct orc_ctx {
  orc_params p;
  int N, Nk, nxh;
  int Mb[3];
  int start[3];     /* first contour index of each block */
  double ds;
  long long* off;   /* slice offsets (doubles), Ns+2 entries */
  fft3 f;
  gtab tab[3][2][3];  /* [order-1][dir (0 fwd,1 bwd)][blk] */
};

int orc_block_of(const orc_ctx* c, int s) {
  /* target s belongs to block b if start[b] < s <= start[b]+Nblk[b]; s = 0 -> block 0.
   * Junction target points take the lower block in both directions
   * (wormlike_MDE_2dcomp_mod.f90:587-591,704 forward; :1335 backward). */
  for (int b = 0; b < c->p.nblk; b++)
    if (s <= c->start[b] + c->p.Nblk[b]) return b;
  return c->p.nblk - 1;
}

int orc_slice_M(const orc_ctx* c, int s) { return c->Mb[orc_block_of(c, s)]; }
long long orc_slice_offset(const orc_ctx* c, int s) { return c->off[s]; }
long long orc_store_size(const orc_ctx* c) { return c->off[c->p.Ns + 1]; }

orc_ctx* orc_create(const orc_params* p) {
  if (p->nblk < 1 || p->nblk > 3) return NULL;
  int tot = 0;
  for (int b = 0; b < p->nblk; b++) {
    if (p->Nblk[b] < 1 || p->L[b] < 0 || p->L[b] > MAXL) return NULL;
    tot += p->Nblk[b];
  }
  if (tot != p->Ns || p->Ns < 3) return NULL;
  orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
  c->p = *p;
  if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
  c->N = p->nx * p->ny * p->nz;
  c->nxh = p->nx / 2 + 1;
  c->Nk = c->nxh * p->ny * p->nz;
  c->ds = 1.0 / p->Ns; /* initialize_mod.f90:140 */
  int st = 0;
  for (int b = 0; b < p->nblk; b++) {
    c->Mb[b] = (p->L[b] + 1) * (p->L[b] + 1);
    c->start[b] = st;
    st += p->Nblk[b];
  }
  c->off = (long long*)malloc(sizeof(long long) * (p->Ns + 2));
  c->off[0] = 0;
  for (int s = 0; s <= p->Ns; s++) c->off[s + 1] = c->off[s] + (long long)orc_slice_M(c, s) * c->N;
  fft3_init(&c->f, p->nx, p->ny, p->nz);
  return c;
}

void orc_destroy(orc_ctx* c) {
  if (!c) return;
  for (int o = 0; o < 3; o++)
    for (int d = 0; d < 2; d++)
      for (int b = 0; b < 3; b++) free(c->tab[o][d][b].lt);
  fft3_free(&c->f);
  free(c->off);
  free(c);
}

static double cdelta_of(int order, int literal) { /* Gmatrix_2dcomp_mod.f90:384-429 */
  if (order == 1) return 1.0;
  if (order == 2) return 1.5;
  return literal ? (double)(11.0f / 6.0f) : 11.0 / 6.0;
}

/* build (once) the G^-1 table for (order, dir, blk), Gmatrix_2dcomp_mod.f90:440-615 */
static gtab* get_tab(orc_ctx* c, int order, int dir, int blk) {
  gtab* t = &c->tab[order - 1][dir][blk];
  if (t->valid) return t;
  int L = c->p.L[blk], M = c->Mb[blk];
  int lt = M * (M + 1) / 2;
  int lit = c->p.literal;
  t->M = M;
  t->lt = (double*)malloc(sizeof(double) * 2 * (size_t)lt * c->Nk);
  double* R = (double*)malloc(sizeof(double) * 3 * M * M);
  orc_rmatrices(L, lit, R, R + M * M, R + 2 * M * M);
  double cd = cdelta_of(order, lit);
  double sigma = dir == 0 ? 1.0 : -1.0;
  double fx = kfactor(c->p.nx, c->p.box[0], lit);
  double fy = kfactor(c->p.ny, c->p.box[1], lit);
  double fz = kfactor(c->p.nz, c->p.box[2], lit);
  double ds = c->ds, lam = c->p.lam[blk];
  int nxh = c->nxh, ny = c->p.ny, nz = c->p.nz;
#pragma omp parallel
  {
    double* g = (double*)malloc(sizeof(double) * 2 * M * M);
#pragma omp for schedule(dynamic, 16)
    for (int k = 0; k < c->Nk; k++) {
      int kxi = k % nxh, kyi = (k / nxh) % ny, kzi = k / (nxh * ny);
      double kx = signed_k(kxi, c->p.nx) * fx;
      double ky = signed_k(kyi, ny) * fy;
      double kz = signed_k(kzi, nz) * fz;
      for (int i = 0; i < M; i++) {
        int l, m; basis_lm(i, &l, &m);
        for (int j = 0; j < M; j++) {
          g[2 * (i * M + j)] = (i == j) ? cd + ds * lam * l * (l + 1) : 0.0;
          g[2 * (i * M + j) + 1] =
              sigma * ds *
              (R[i * M + j] * kx + R[M * M + i * M + j] * ky + R[2 * M * M + i * M + j] * kz);
        }
      }
      zinv(M, g);
      double* dst = t->lt + (size_t)2 * lt * k;
      int e = 0;
      for (int i = 0; i < M; i++)
        for (int j = 0; j <= i; j++, e++) {
          dst[2 * e] = g[2 * (i * M + j)];
          dst[2 * e + 1] = g[2 * (i * M + j) + 1];
        }
    }
    free(g);
  }
  free(R);
  t->valid = 1;
  return t;
}

/* the RHS rule of wormlike_MDE_2dcomp_mod.f90:325-343,484-502,584-608 (and the
 * factored copies in init_pr_forward.f90 / init_pr_backward.f90) */
static void build_rhs(const orc_ctx* c, int order, const double* w, const double* const h[3],
                      const int hM[3], int M, double* P) {
  int N = c->N;
  double ds = c->ds;
  double third = c->p.literal ? (double)(1.0f / 3.0f) : 1.0 / 3.0;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < M; i++) {
    const double* q1 = (h[0] && i < hM[0]) ? h[0] + (size_t)i * N : NULL;
    const double* q2 = (order >= 2 && h[1] && i < hM[1]) ? h[1] + (size_t)i * N : NULL;
    const double* q3 = (order >= 3 && h[2] && i < hM[2]) ? h[2] + (size_t)i * N : NULL;
    double* out = P + (size_t)i * N;
    for (int r = 0; r < N; r++) {
      double a1 = q1 ? q1[r] : 0.0, a2 = q2 ? q2[r] : 0.0, a3 = q3 ? q3[r] : 0.0;
      double W = w[r];
      double v;
      if (order == 1) {
        v = 0.0 + (1.0 - ds * W) * a1;
      } else if (order == 2) {
        v = 0.0 + 2.0 * a1 - 0.5 * a2 - ds * W * (2.0 * a1 - a2);
      } else {
        v = 0.0 + 3.0 * a1 - 1.5 * a2 + third * a3 - ds * W * (3.0 * a1 - 3.0 * a2 + a3);
      }
      out[r] = v;
    }
  }
}

/* FFT -> G^-1 (lower triangle, symmetric) -> iFFT / N : wormlike_MDE_2dcomp_mod.f90:357-455 */
static void solve_slice(orc_ctx* c, gtab* t, int M, const double* P, double* qout) {
  int N = c->N, Nk = c->Nk;
  double* spec = (double*)malloc(sizeof(double) * 2 * (size_t)Nk * M);
  for (int i = 0; i < M; i++) fft3_r2c(&c->f, P + (size_t)i * N, spec + (size_t)2 * i * Nk);
  int lt = M * (M + 1) / 2;
  int lit = c->p.literal;
#pragma omp parallel
  {
    double* pr = (double*)malloc(sizeof(double) * 4 * M);
    double* pi = pr + M;
    double* yr = pi + M;
    double* yi = yr + M;
#pragma omp for schedule(static)
    for (int k = 0; k < Nk; k++) {
      for (int i = 0; i < M; i++) {
        pr[i] = spec[2 * ((size_t)i * Nk + k)];
        pi[i] = spec[2 * ((size_t)i * Nk + k) + 1];
      }
      const double* A = t->lt + (size_t)2 * lt * k;
      /* 4 x dcsrsymv('l'): RR = R Pr, RI = R Pi, IR = I Pr, II = I Pi */
      for (int i = 0; i < M; i++) {
        double RR = 0, RI = 0, IR = 0, II = 0;
        for (int j = 0; j < M; j++) {
          int e = (j <= i) ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
          double ar = A[2 * e], ai = A[2 * e + 1];
          RR += ar * pr[j]; RI += ar * pi[j];
          IR += ai * pr[j]; II += ai * pi[j];
        }
        yr[i] = RR - II;
        yi[i] = RI + IR;
        if (lit) { /* cmplx(...) of default kind, :429,442 */
          yr[i] = (double)(float)yr[i];
          yi[i] = (double)(float)yi[i];
        }
      }
      for (int i = 0; i < M; i++) {
        spec[2 * ((size_t)i * Nk + k)] = yr[i];
        spec[2 * ((size_t)i * Nk + k) + 1] = yi[i];
      }
    }
    free(pr);
  }
  double inv = 1.0 / (double)N; /* M_grid_inv = 1.0d0/M_grid */
  for (int i = 0; i < M; i++) {
    double* o = qout + (size_t)i * N;
    fft3_c2r(&c->f, spec + (size_t)2 * i * Nk, o);
    for (int r = 0; r < N; r++) o[r] *= inv;
  }
  free(spec);
}

int orc_step(orc_ctx* c, const double* w, int dir, int s, const double* const hist[3],
             double* out) {
  int Ns = c->p.Ns;
  int order = dir > 0 ? (s >= 3 ? 3 : s) : (Ns - s >= 3 ? 3 : Ns - s);
  if (order < 1) return -1;
  int blk = orc_block_of(c, s);
  int M = c->Mb[blk];
  int hM[3];
  for (int j = 0; j < 3; j++) {
    int sh = s - dir * (j + 1);
    hM[j] = (sh >= 0 && sh <= Ns) ? orc_slice_M(c, sh) : 0;
  }
  double* P = (double*)malloc(sizeof(double) * (size_t)M * c->N);
  build_rhs(c, order, w + (size_t)blk * c->N, hist, hM, M, P);
  gtab* t = get_tab(c, order, dir > 0 ? 0 : 1, blk);
  solve_slice(c, t, M, P, out);
  free(P);
  return 0;
}

int orc_forward(orc_ctx* c, const double* w, double* q) {
  int N = c->N, Ns = c->p.Ns;
  int M0 = orc_slice_M(c, 0);
  memset(q, 0, sizeof(double) * (size_t)M0 * N); /* qA(0) = delta_{i0}, :309-314 */
  for (int r = 0; r < N; r++) q[r] = 1.0;
  for (int s = 1; s <= Ns; s++) {
    const double* h[3];
    for (int j = 0; j < 3; j++) h[j] = (s - j - 1 >= 0) ? q + c->off[s - j - 1] : NULL;
    if (orc_step(c, w, +1, s, h, q + c->off[s])) return -1;
  }
  return 0;
}

int orc_backward(orc_ctx* c, const double* w, double* qs) {
  int N = c->N, Ns = c->p.Ns;
  int Me = orc_slice_M(c, Ns);
  double* last = qs + c->off[Ns];
  memset(last, 0, sizeof(double) * (size_t)Me * N); /* qBstar(NB) = delta, :1035-1040 */
  for (int r = 0; r < N; r++) last[r] = 1.0;
  for (int s = Ns - 1; s >= 0; s--) {
    const double* h[3];
    for (int j = 0; j < 3; j++) h[j] = (s + j + 1 <= Ns) ? qs + c->off[s + j + 1] : NULL;
    if (orc_step(c, w, -1, s, h, qs + c->off[s])) return -1;
  }
  return 0;
}

int orc_density(orc_ctx* c, const double* q, const double* qs, double* phi, double* Q,
                double* totden) {
  int N = c->N, nb = c->p.nblk;
  double V = c->p.box[0] * c->p.box[1] * c->p.box[2];
  /* Q = (1/V) int q_0(r, Ns) dr, SCFT_Density_FE.f90:48-91 */
  *Q = orc_volume_integral(c->p.nx, c->p.ny, c->p.nz, c->p.box, q + c->off[c->p.Ns]) / V;
  double* denz = (double*)calloc((size_t)N, sizeof(double));
  int maxn = 0;
  for (int b = 0; b < nb; b++) if (c->p.Nblk[b] > maxn) maxn = c->p.Nblk[b];
#pragma omp parallel
  {
    double* a = (double*)malloc(sizeof(double) * (maxn + 1));
#pragma omp for schedule(static)
    for (int r = 0; r < N; r++) {
      double tot = 0.0;
      for (int b = 0; b < nb; b++) {
        int n = c->p.Nblk[b], s0 = c->start[b];
        int Mb = c->Mb[b];
        double sum_i = 0.0;
        for (int i = 0; i < Mb; i++) {
          for (int j = 0; j <= n; j++) {
            int s = s0 + j;
            int Ms = orc_slice_M(c, s);
            /* block b's first point is the previous block's junction slice; only the
             * harmonics both blocks share enter (qB(0)=qA(NA) truncated, :927-937) */
            int lim = Ms < Mb ? Ms : Mb;
            a[j] = (i < lim) ? q[c->off[s] + (size_t)i * N + r] * qs[c->off[s] + (size_t)i * N + r]
                             : 0.0;
          }
          sum_i += simpson_1d(n, c->ds, a); /* RHO_i = simposon_1D_NR(0,N_blk,ds,c) */
        }
        phi[(size_t)b * N + r] = sum_i;
        tot += sum_i;
      }
      denz[r] = tot;
    }
    free(a);
  }
  double td = orc_volume_integral(c->p.nx, c->p.ny, c->p.nz, c->p.box, denz);
  *totden = td;
  for (int b = 0; b < nb; b++)
    for (int r = 0; r < N; r++) phi[(size_t)b * N + r] = V * phi[(size_t)b * N + r] / td;
  free(denz);
  return 0;
}

double orc_free_energy(const orc_ctx* c, const double* w, const double* phi, double Q,
                       const double* chi) {
  int N = c->N, nb = c->p.nblk;
  double V = c->p.box[0] * c->p.box[1] * c->p.box[2];
  double* ar = (double*)malloc(sizeof(double) * N);
  for (int r = 0; r < N; r++) {
    double xi = 0.0, sphi = 0.0, e = 0.0;
    for (int b = 0; b < nb; b++) { xi += w[(size_t)b * N + r]; sphi += phi[(size_t)b * N + r]; }
    xi /= nb;
    for (int a = 0; a < nb; a++)
      for (int b = a + 1; b < nb; b++) e += chi[a * 3 + b] * phi[(size_t)a * N + r] * phi[(size_t)b * N + r];
    for (int b = 0; b < nb; b++) e -= w[(size_t)b * N + r] * phi[(size_t)b * N + r];
    e -= xi * (1.0 - sphi);
    ar[r] = e;
  }
  double I = orc_volume_integral(c->p.nx, c->p.ny, c->p.nz, c->p.box, ar) / V;
  free(ar);
  return -log(Q) + I;
}

void orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}

double orc_time_bdf3_steps(orc_ctx* c, const double* w, int nsteps) {
  int N = c->N;
  int M = c->Mb[0];
  gtab* t = get_tab(c, 3, 0, 0); /* factor build, excluded from the timing */
  double* h = (double*)malloc(sizeof(double) * 4 * (size_t)M * N);
  for (size_t e = 0; e < 4 * (size_t)M * N; e++) h[e] = 0.0;
  for (int r = 0; r < N; r++) { h[r] = 1.0; h[(size_t)M * N + r] = 1.0; h[(size_t)2 * M * N + r] = 1.0; }
  double* P = (double*)malloc(sizeof(double) * (size_t)M * N);
  int hM[3] = {M, M, M};
  double t0 = omp_get_wtime();
  for (int it = 0; it < nsteps; it++) {
    int a = it % 4;
    const double* hh[3] = {h + (size_t)((a + 3) % 4) * M * N, h + (size_t)((a + 2) % 4) * M * N,
                           h + (size_t)((a + 1) % 4) * M * N};
    build_rhs(c, 3, w, hh, hM, M, P);
    solve_slice(c, t, M, P, h + (size_t)a * M * N);
  }
  double t1 = omp_get_wtime();
  free(P);
  free(h);
  return t1 - t0;
}
