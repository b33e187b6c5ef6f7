// Host dispatch of the per-Lmax kernel-B translation units (wlc_zsolve.hip is
// compiled once per L = 0..8) and of the orientation-table uploads.
#include <hip/hip_runtime.h>

#include "wlc_kernels.h"

namespace wlc {

#define WLC_DECL(L)                                                                          \
  hipError_t zsolve_launch_L##L(const Geom&, const SolveArgs&, hipStream_t);                 \
  hipError_t zsolve_batch_L##L(const ZJob&, const ZJob*, int, hipStream_t);                 \
  LaunchKey zsolve_key_L##L(const Geom&, const SolveArgs&);                                  \
  hipError_t test_solve_L##L(int, const double*, double, double, double, double, double2*,   \
                             hipStream_t);                                                   \

WLC_DECL(0) WLC_DECL(1) WLC_DECL(2) WLC_DECL(3) WLC_DECL(4) WLC_DECL(5) WLC_DECL(6) WLC_DECL(7) WLC_DECL(8)
#undef WLC_DECL

hipError_t launch_zsolve(const Geom& g, const SolveArgs& a, hipStream_t st) {
  switch (a.L) {
#define WLC_CASE(L) \
  case L: return zsolve_launch_L##L(g, a, st);
    WLC_CASE(0) WLC_CASE(1) WLC_CASE(2) WLC_CASE(3) WLC_CASE(4) WLC_CASE(5) WLC_CASE(6) WLC_CASE(7) WLC_CASE(8)
#undef WLC_CASE
  }
  return hipErrorInvalidValue;
}

hipError_t launch_zsolve_batch(const ZJob& rep, const ZJob* jobs, int n, hipStream_t st) {
  switch (rep.a.L) {
#define WLC_CASE(L) \
  case L: return zsolve_batch_L##L(rep, jobs, n, st);
    WLC_CASE(0) WLC_CASE(1) WLC_CASE(2) WLC_CASE(3) WLC_CASE(4) WLC_CASE(5) WLC_CASE(6) WLC_CASE(7) WLC_CASE(8)
#undef WLC_CASE
  }
  return hipErrorInvalidValue;
}

LaunchKey zsolve_key(const Geom& g, const SolveArgs& a) {
  switch (a.L) {
#define WLC_CASE(L) \
  case L: return zsolve_key_L##L(g, a);
    WLC_CASE(0) WLC_CASE(1) WLC_CASE(2) WLC_CASE(3) WLC_CASE(4) WLC_CASE(5) WLC_CASE(6) WLC_CASE(7) WLC_CASE(8)
#undef WLC_CASE
  }
  return LaunchKey{nullptr, 0, 0, 0, 0};
}

hipError_t launch_test_solve(int L, int nk, const double* kvec, double cdelta, double sigma,
                             double ds, double dslam, double2* v, hipStream_t st) {
  switch (L) {
#define WLC_CASE(LL) \
  case LL: return test_solve_L##LL(nk, kvec, cdelta, sigma, ds, dslam, v, st);
    WLC_CASE(0) WLC_CASE(1) WLC_CASE(2) WLC_CASE(3) WLC_CASE(4) WLC_CASE(5) WLC_CASE(6) WLC_CASE(7) WLC_CASE(8)
#undef WLC_CASE
  }
  return hipErrorInvalidValue;
}

namespace {
#define WLC_TABLE_QUAL static constexpr
#include "wlc_tables.inc"
#undef WLC_TABLE_QUAL
}  // namespace

// The device code carries the orientation tables as compile-time constants; check
// them against the ones the host set-up code derives now.
hipError_t upload_orientation_tables(const OrientTables& t) {
  for (int i = 0; i < kMaxM; ++i)
    if (kZsign[i] != t.zsign[i] || kTri[i] != t.tri_e[i]) return hipErrorInvalidValue;
  const int ny = t.yh_off[kMaxL + 1];
  if (ny != (int)(sizeof(kYh) / sizeof(kYh[0]))) return hipErrorInvalidValue;
  for (int i = 0; i < ny; ++i)
    if (kYh[i] != t.yh[i]) return hipErrorInvalidValue;
  return hipSuccess;
}

}  // namespace wlc
