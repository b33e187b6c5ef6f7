// Host helpers shared by the handle (wlc_mde.hip) and the batch (wlc_batch.hip) of
// libwlc_mde.so (internal, not part of the C ABI).
#pragma once
#include <string>
#include <vector>

#include "../../include/wlc_mde.h"

namespace wlc {
// records msg as wlc_mde_last_error() of this thread and returns code
int host_fail(int code, const std::string& msg);
// the checks of wlc_mde_create (grid, blocks, Lmax, lambda, plane size)
int validate_params(const wlc_params* p);
// contour weights of a block of n intervals: composite Simpson (n even: the reference's
// rule, utility_mod.f90:346-374), odd n closes with a 3/8 panel, n = 1 trapezoid
void simpson_weights(int n, double ds, std::vector<double>* w);
// kx columns per workgroup of the generic z solve
int zsolve_tx(int nz);
// kernel-B variant: the autotuned table entry for (nz, L, spectrum precision) or the
// built-in heuristic, with the measurement environment switches applied on top
struct ZOpts {
  int generic, split, occ, tx;
  int tuned;  // 1: from the autotuned table
};
ZOpts zsolve_opts(int nz, int L, int spec32);
// orientation tables built and checked for this device
int orientation_tables_ready(int device);
// BDF1/2/3 right-hand-side coefficients and the diagonal c_delta of G
void bdf(int order, double a[3], double b[3], double* cdelta);
}  // namespace wlc
