// Internal declarations shared by the host set-up code and the HIP kernels of
// libwlc_mde.so (not part of the C ABI).
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace wlc {

constexpr int kMaxL = 8;                            // largest Lmax a block may use
constexpr int kMaxM = (kMaxL + 1) * (kMaxL + 1);    // 81 harmonics
constexpr int kMaxBlocks = 3;

inline int sh_index(int l, int m) { return l * l + l + m; }

// Orientation tables for the k-aligned solve, built once per handle on the host
// and copied to __constant__ memory (identical for every block: they depend only
// on the basis convention, truncated at the block's own Lmax).
struct OrientTables {
  // Zr(alpha) = rep. of the rotation about z by alpha: element (i, partner(i))
  // equals zsign[i] * sin(|m| alpha); diagonal is cos(|m| alpha).
  double zsign[kMaxM];
  // Yh = rep. of the rotation about y by pi/2, packed: for l = 0..kMaxL, rows m
  // in -l..l, columns m' in -l..l restricted to the nonzero pattern
  // (same sign class of m and m', and (m + m' + l) even for m >= 0 / odd for m < 0).
  double yh[4096];
  int yh_off[kMaxL + 2];
  // tridiagonal coupling of the rotated frame: e[i] = Rz[(l,m),(l+1,m)]
  double tri_e[kMaxM];
};

// true when (l, m, m') is in the structural nonzero pattern of Yh
inline bool yh_pattern(int l, int m, int mp) {
  if ((m >= 0) != (mp >= 0)) return false;
  int par = (m + mp + l) & 1;
  return (m >= 0) ? (par == 0) : (par == 1);
}

// Host set-up: builds R matrices (real-SH Gaunt via the complex Gaunt coefficient
// and a unitary real<->complex map), the rotation tables, and self-checks the
// rotated-frame solve against a dense complex solve.  Returns "" on success.
std::string build_orientation_tables(int Lmax, OrientTables* t,
                                     std::vector<double>* Rx = nullptr,
                                     std::vector<double>* Ry = nullptr,
                                     std::vector<double>* Rz = nullptr);

// Host reference of the rotated-frame solve (same arithmetic as the kernel); used
// by the set-up self-check and by the test entry point.
void host_rotated_solve(const OrientTables& t, int L, double cdelta, double sigma, double ds,
                        double dslam, const double k[3], double* v /* M complex interleaved */);

// FFT plan along one axis: radices (8, 4, 3, 2) and a twiddle table exp(-2 pi i t/n).
struct AxisPlan {
  int n = 0;
  int npass = 0;
  int radix[16] = {0};
  std::vector<double> tw;  // 2n doubles
};
std::string make_axis_plan(int n, AxisPlan* p);

}  // namespace wlc
