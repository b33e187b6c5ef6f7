// Kernel-B (z FFT + solve) variant per (nz, Lmax, spectrum precision): the table written by
// the autotuner (autotune.py: every admissible variant timed on the device, the fastest
// kept), else the built-in heuristic; the measurement switches WLC_GENERIC_ZSOLVE,
// WLC_ZSOLVE_SPLIT, WLC_ZSOLVE_OCC and WLC_ZSOLVE_TX override either.
#include <cstdlib>

#include "wlc_host.h"

namespace wlc {
namespace {

struct TuneRow {
  int nz, L, spec32;
  int generic, split, occ, tx;
  double us;  // measured kernel-B time of the chosen variant (documentation only)
};

const TuneRow kTune[] = {
#include "wlc_tune.inc"
    {0, 0, 0, 0, 0, 0, 0, 0.0}};

int env_int(const char* name, int fallback) {
  const char* v = std::getenv(name);
  return v ? std::atoi(v) : fallback;
}

}  // namespace

ZOpts zsolve_opts(int nz, int L, int spec32) {
  ZOpts o;
  o.generic = 0;
  o.occ = 1;
  o.split = 2;  // Lmax 5-7 at nz = 64: the whole-tile split kernel; elsewhere a degree split
  if ((nz == 24 || nz == 32 || nz == 48) && L >= 4 && L <= 7 && !spec32)
    o.split = 3;  // the half-tile kernel with a wave-local transform (measured faster at every
                  // tabulated Lmax of these edges)
  o.tx = zsolve_tx(nz);
  o.tuned = 0;
  for (const TuneRow& r : kTune) {
    if (r.nz == nz && r.L == L && r.spec32 == spec32) {
      o.generic = r.generic;
      o.split = r.split;
      o.occ = r.occ;
      o.tx = r.tx;
      o.tuned = 1;
      break;
    }
  }
  o.generic = env_int("WLC_GENERIC_ZSOLVE", o.generic) != 0;
  {
    const int v = env_int("WLC_ZSOLVE_OCC", o.occ);
    o.occ = v == 2 ? 2 : 1;
  }
  {
    const int v = env_int("WLC_ZSOLVE_SPLIT", o.split);
    o.split = (v >= 1 && v <= 3) ? v : 0;
  }
  {
    const int v = env_int("WLC_ZSOLVE_TX", o.tx);
    if (v >= 1 && v * nz <= 256) o.tx = v;
  }
  return o;
}

}  // namespace wlc
