// C ABI of libwlc_mde.so: handle, propagator stores and the contour sweeps.
//
// Forward sweep (reference MDE_q, wormlike_MDE_2dcomp_mod.f90:231-956):
//   q(0) = delta_{i0}; for s = 1..Ns: BDF order min(s, 3); the target point s takes
//   the field, Lmax and lambda of the block that contains it (junction point = lower
//   block, :587-591, :704); history slices of another block are zero-padded or
//   truncated to the target block's harmonics (:727-734).
// Backward sweep (reference MDE_qstar, :961-1557): mirror with sigma = -1, BDF order
//   min(Ns - s, 3), q*(Ns) = delta_{i0}; while it runs, the per-block contour Simpson
//   sum of sum_i q_i q*_i is accumulated (density, SCFT_Density_FE.f90:97-118).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/wlc_mde.h"
#include "wlc_host.h"
#include "wlc_internal.h"
#include "wlc_kernels.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIPCHK(call)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(WLC_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));          \
  } while (0)

std::mutex g_tab_mu;
bool g_tab_built = false;
wlc::OrientTables g_tab;
std::vector<int> g_tab_uploaded;  // per device

// composite Simpson over n intervals (n even: the reference's rule,
// utility_mod.f90:346-374); odd n >= 3 closes with a 3/8 panel; n = 1 trapezoid
void contour_weights(int n, double ds, std::vector<double>* w) {
  w->assign(n + 1, 0.0);
  if (n == 1) {
    (*w)[0] = (*w)[1] = ds / 2.0;
    return;
  }
  const int ns = (n % 2 == 0) ? n : n - 3;
  for (int j = 0; j <= ns && ns > 0; ++j) {
    const double c = (j == 0 || j == ns) ? 1.0 : ((j % 2) ? 4.0 : 2.0);
    (*w)[j] += (ds / 3.0) * c;
  }
  if (n % 2) {
    const double c = 3.0 * ds / 8.0;
    (*w)[ns] += c;
    (*w)[ns + 1] += 3.0 * c;
    (*w)[ns + 2] += 3.0 * c;
    (*w)[ns + 3] += c;
  }
}

}  // namespace

struct wlc_handle {
  wlc_params p{};
  int device = 0;
  long long N = 0, Nk = 0;
  long long zs = 0, hs = 0;  // padded real-space strides
  int nxh = 0, Ns = 0, Mmax = 0;
  double ds = 0.0;
  int Mb[3] = {0, 0, 0}, Lb[3] = {0, 0, 0}, start[3] = {0, 0, 0};
  std::vector<long long> off;  // forward store slice offsets (doubles)
  std::vector<std::vector<double>> cw;  // contour weights per block
  double* q = nullptr;
  double* qs = nullptr;        // backward: full store (keep_qstar) or 4-slot ring
  double2* spec = nullptr;
  double* w = nullptr;
  double* zero = nullptr;  // one field of zeros: the history plane of a harmonic a slice lacks
  double* rho = nullptr;
  double* phi = nullptr;
  double* partial = nullptr;
  double* scal = nullptr;      // [0] = int q_0(Ns), [1] = totden
  int nparts = 256;
  double2* tw[3] = {nullptr, nullptr, nullptr};
  wlc::AxisPlan plan[3];
  wlc::Geom geom{};
  hipStream_t st = nullptr;
  // The backward sweep runs the density accumulation of slice s on st2, overlapped
  // with the z solve of the next step on st: st2 waits for ev_step (slice s complete),
  // and before the backward ring slot of slice s is rewritten (four steps later) st
  // waits for ev_dens[s & 3].
  hipStream_t st2 = nullptr;
  hipEvent_t ev_step = nullptr;
  hipEvent_t ev_dens[4] = {nullptr, nullptr, nullptr, nullptr};
  bool dens_used[4] = {false, false, false, false};
  // Concurrent sweeps (when the full backward store fits): the backward sweep runs on
  // stb with its own half spectrum specb while the forward sweep runs on st; slice s of
  // either sweep records evf[s] / evb[s].  Density term s runs once both slices exist:
  // on st2 into rho for s <= Ns/2 (ready as the backward sweep reaches s), on st3 into
  // rho2 for s > Ns/2 (ready as the forward sweep reaches s).  The backward sweep starts
  // after ev_pre, recorded on st before the forward sweep is queued; every stream is
  // joined into st before wlc_mde_backward returns.
  bool conc = false;
  hipStream_t stb = nullptr, st3 = nullptr;
  double2* specb = nullptr;
  double* rho2 = nullptr;
  std::vector<hipEvent_t> evf, evb;
  hipEvent_t ev_pre = nullptr, ev_join[3] = {nullptr, nullptr, nullptr};
  bool pre_recorded = false;
  // Forward-store checkpointing (ck > 0, when the full store does not fit or on request):
  // the forward sweep keeps a 4-slot ring `fr` and copies slices jK-2, jK-1, jK of every
  // segment boundary jK (j = 1 .. nseg-1) to `ckpt` and q(Ns) to `qlast`.  The backward
  // sweep recomputes segment j = (jK, (j+1)K] (segment 0 = [0, K]) from its checkpoint into
  // segment buffer j % 2 on stream `str` with its own spectrum `specr`, one segment ahead
  // of the densities that read it: ev_seg_ready[b] marks a buffer filled, ev_seg_free[b]
  // (on st2, after the last density of the segment it held) marks it reusable.  The
  // recomputed slices are bit-identical to the first forward sweep's.
  int ck = 0, nseg = 0;
  // Hybrid store: the lowest nstored segments (slices 0 .. nstored K) are kept whole in
  // `store` by the forward sweep and read in place by the backward densities; only the
  // segments above are checkpointed and recomputed.
  int nstored = 0;
  double* store = nullptr;
  double* fr = nullptr;
  double* ckpt = nullptr;
  double* segbuf[2] = {nullptr, nullptr};
  double* qlast = nullptr;
  double2* specr = nullptr;
  hipStream_t str = nullptr;
  hipEvent_t ev_seg_ready[2] = {nullptr, nullptr}, ev_seg_free[2] = {nullptr, nullptr};
  bool seg_busy[2] = {false, false};
  int dens_seg = -1;  // segment whose densities are being queued
  bool have_w = false, fwd_done = false, bwd_done = false;
  int fuse = 1;  // kernel C of a step fused with kernel A of the next
  // timing
  bool timing = false;
  bool timing_serial = false;  // timing mode 2: the backward sweep waits for the forward
  struct Rec { int k; hipEvent_t a, b; };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  double ms[5] = {0, 0, 0, 0, 0};
  long long cnt[5] = {0, 0, 0, 0, 0};
  hipEvent_t uev[8] = {};
};

namespace {

// where forward slice s of the sweep being queued lives
struct FwdLoc {
  int mode = 0;  // 0: full store, 1: 4-slot ring, 2: segment buffer starting at slice lo
  double* base = nullptr;
  int lo = 0;
};

// one sweep being queued: direction, stream, half spectrum, forward-slice locator
struct SweepCtx {
  int dir;
  hipStream_t st;
  double2* spec;
  FwdLoc fl;
};

long long slot_size(const wlc_handle* h) { return (long long)h->Mmax * h->hs; }

double* fwd_slice(const wlc_handle* h, const FwdLoc& fl, int s) {
  if (fl.mode == 1) {
    if (h->nstored > 0 && s <= h->nstored * h->ck) return h->store + (long long)s * slot_size(h);
    return h->fr + (long long)(s & 3) * slot_size(h);
  }
  if (fl.mode == 2) return fl.base + (long long)(s - fl.lo) * slot_size(h);
  return h->q + h->off[s];
}

double* fwd_last(const wlc_handle* h) { return h->ck ? h->qlast : h->q + h->off[h->Ns]; }

int block_of(const wlc_handle* h, int s) {
  for (int b = 0; b < h->p.nblk; ++b)
    if (s <= h->start[b] + h->p.Nblk[b]) return b;
  return h->p.nblk - 1;
}

double* qs_slot(wlc_handle* h, int s) {
  if (h->p.keep_qstar || h->conc) return h->qs + h->off[s];
  return h->qs + (long long)(s & 3) * h->Mmax * h->hs;
}

hipEvent_t take_event(wlc_handle* h) {
  if (!h->pool.empty()) {
    hipEvent_t e = h->pool.back();
    h->pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int flush_timing(wlc_handle* h) {
  if (h->recs.empty()) return WLC_OK;
  HIPCHK(hipStreamSynchronize(h->st));
  HIPCHK(hipStreamSynchronize(h->st2));
  if (h->stb) HIPCHK(hipStreamSynchronize(h->stb));
  if (h->st3) HIPCHK(hipStreamSynchronize(h->st3));
  for (auto& r : h->recs) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
    h->ms[r.k] += t;
    h->cnt[r.k] += 1;
    h->pool.push_back(r.a);
    h->pool.push_back(r.b);
  }
  h->recs.clear();
  return WLC_OK;
}

struct TimedLaunch {
  wlc_handle* h;
  int k;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  TimedLaunch(wlc_handle* hh, int kk, hipStream_t ss = nullptr) : h(hh), k(kk), s(ss ? ss : hh->st) {
    if (h->timing) {
      a = take_event(h);
      b = take_event(h);
      if (a) (void)hipEventRecord(a, s);
    }
  }
  ~TimedLaunch() {
    if (h->timing && a && b) {
      (void)hipEventRecord(b, s);
      h->recs.push_back({k, a, b});
    }
  }
};

// dense [nf][nz][ny][nx] <-> padded device layout (plane stride zs, field stride hs)
hipError_t copy_fields_in(wlc_handle* h, double* dst, const double* src, int nf, hipMemcpyKind kind) {
  const size_t row = sizeof(double) * (size_t)h->p.nx * h->p.ny;
  for (int f = 0; f < nf; ++f) {
    hipError_t e = hipMemcpy2DAsync(dst + (long long)f * h->hs, sizeof(double) * h->zs,
                                    src + (long long)f * h->N, row, row, h->p.nz, kind, h->st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t copy_fields_out(wlc_handle* h, double* dst, const double* src, int nf, hipMemcpyKind kind) {
  const size_t row = sizeof(double) * (size_t)h->p.nx * h->p.ny;
  for (int f = 0; f < nf; ++f) {
    hipError_t e = hipMemcpy2DAsync(dst + (long long)f * h->N, row, src + (long long)f * h->hs,
                                    sizeof(double) * h->zs, row, h->p.nz, kind, h->st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// the main sweep of direction dir (forward: full store or ring; backward: own stream and
// spectrum when the sweeps run concurrently)
SweepCtx main_sweep(const wlc_handle* h, int dir) {
  SweepCtx c{dir, h->st, h->spec, FwdLoc{}};
  if (dir < 0 && h->conc) {
    c.st = h->stb;
    c.spec = h->specb;
  }
  if (h->ck) c.fl.mode = 1;
  return c;
}

// BDF1/2/3 coefficients: P = sum_j a_j q(s - dir (j + 1)) - ds w sum_j b_j q(...), and the
// diagonal c_delta of G (wormlike_MDE_2dcomp_mod.f90:341, :499-500, :604-606;
// Gmatrix_2dcomp_mod.f90:384-429)
void bdf_coeffs(int order, double a[3], double b[3], double* cdelta) {
  if (order == 1) {
    a[0] = 1.0; a[1] = 0.0; a[2] = 0.0;
    b[0] = 1.0; b[1] = 0.0; b[2] = 0.0;
    *cdelta = 1.0;
  } else if (order == 2) {
    a[0] = 2.0; a[1] = -0.5; a[2] = 0.0;
    b[0] = 2.0; b[1] = -1.0; b[2] = 0.0;
    *cdelta = 1.5;
  } else {
    a[0] = 3.0; a[1] = -1.5; a[2] = 1.0 / 3.0;
    b[0] = 3.0; b[1] = -3.0; b[2] = 1.0;
    *cdelta = 11.0 / 6.0;
  }
}

// Right-hand-side arguments of the step whose target is s.  BDF order by distance from
// the chain end the sweep starts at (wormlike_MDE_2dcomp_mod.f90:325-343 BDF1, :484-502
// BDF2, :584-608 BDF3; backward :1046-1064, :1138-1157, :1234-1259); the field is the
// target block's; history j is slice s - dir (j + 1) with its own harmonic count.
void rhs_args(wlc_handle* h, const SweepCtx& sw, int s, wlc::RhsArgs* ra, int* order,
              double* cdelta) {
  const int Ns = h->Ns;
  const int dir = sw.dir;
  *order = dir > 0 ? std::min(s, 3) : std::min(Ns - s, 3);
  const int blk = block_of(h, s);
  for (int j = 0; j < 3; ++j) {
    const int sh = s - dir * (j + 1);
    if (j < *order && sh >= 0 && sh <= Ns) {
      ra->h[j] = dir > 0 ? fwd_slice(h, sw.fl, sh) : qs_slot(h, sh);
      ra->hM[j] = h->Mb[block_of(h, sh)];
    } else {
      ra->h[j] = nullptr;
      ra->hM[j] = 0;
    }
  }
  bdf_coeffs(*order, ra->a, ra->b, cdelta);
  ra->ds = h->ds;
  ra->w = h->w + (long long)blk * h->hs;
  ra->zero = h->zero;
  ra->M = h->Mb[blk];
  ra->spec = sw.spec;
}

double* slice_of(wlc_handle* h, const SweepCtx& sw, int s) {
  return sw.dir > 0 ? fwd_slice(h, sw.fl, s) : qs_slot(h, s);
}

// kernel A for target s (first step of a sweep, or every step when unfused)
int step_rhs(wlc_handle* h, const SweepCtx& sw, int s) {
  wlc::RhsArgs ra{};
  int order;
  double cdelta;
  rhs_args(h, sw, s, &ra, &order, &cdelta);
  TimedLaunch t(h, 0, sw.st);
  HIPCHK(wlc::launch_rhs_fft(h->geom, ra, order, sw.st));
  return WLC_OK;
}

// kernel B for target s: z pass, solve of G^{c,sigma}(k) (Gmatrix_2dcomp_mod.f90:475-479),
// inverse z pass
int step_solve(wlc_handle* h, const SweepCtx& sw, int s) {
  const int Ns = h->Ns;
  const int dir = sw.dir;
  const int order = dir > 0 ? std::min(s, 3) : std::min(Ns - s, 3);
  const int blk = block_of(h, s);
  wlc::SolveArgs sa{};
  sa.spec = sw.spec;
  sa.M = h->Mb[blk];
  sa.L = h->Lb[blk];
  const wlc::ZOpts zo = wlc::zsolve_opts(h->p.nz, h->Lb[blk], h->p.spectrum_fp32);
  sa.TX = zo.tx;
  sa.cdelta = order == 1 ? 1.0 : (order == 2 ? 1.5 : 11.0 / 6.0);
  sa.sigma = dir > 0 ? 1.0 : -1.0;
  sa.ds = h->ds;
  sa.dslam = h->ds * h->p.lambda[blk];
  const double tp = 2.0 * M_PI;
  for (int a = 0; a < 3; ++a) sa.kfac[a] = tp / h->p.box[a];
  sa.solve = 1;
  sa.generic = zo.generic;
  sa.occ = zo.occ;
  sa.split = zo.split;
  TimedLaunch t(h, 1, sw.st);
  HIPCHK(wlc::launch_zsolve(h->geom, sa, sw.st));
  return WLC_OK;
}

// kernel C for target s (inverse xy, 1/N, store), fused with kernel A of the next
// target s + dir when `next` is set
int step_inverse(wlc_handle* h, const SweepCtx& sw, int s, bool next) {
  const int dir = sw.dir;
  const hipStream_t ss = sw.st;
  if (dir < 0 && h->dens_used[s & 3]) HIPCHK(hipStreamWaitEvent(ss, h->ev_dens[s & 3], 0));
  wlc::InvArgs ia{};
  ia.spec = sw.spec;
  ia.M = h->Mb[block_of(h, s)];
  ia.scale = 1.0 / (double)h->N;
  ia.out = slice_of(h, sw, s);
  if (!next) {
    TimedLaunch t(h, 2, ss);
    HIPCHK(wlc::launch_ifft_store(h->geom, ia, ss));
  } else {
    wlc::RhsArgs ra{};
    int order;
    double cdelta;
    rhs_args(h, sw, s + dir, &ra, &order, &cdelta);
    TimedLaunch t(h, 4, ss);
    HIPCHK(wlc::launch_inv_rhs(h->geom, ia, ra, order, ss));
  }
  if (h->conc && sw.fl.mode == 0) HIPCHK(hipEventRecord(dir > 0 ? h->evf[s] : h->evb[s], ss));
  return WLC_OK;
}

int density_at(wlc_handle* h, int s);

// the steps of one sweep, targets first .. last (inclusive, in direction dir); after
// each target the optional density accumulation
int save_checkpoint(wlc_handle* h, const SweepCtx& sw, int s);

int sweep(wlc_handle* h, const SweepCtx& sw, int first, int last, bool dens) {
  const int dir = sw.dir;
  int rc = step_rhs(h, sw, first);
  if (rc) return rc;
  for (int s = first;; s += dir) {
    const bool more = s != last;
    if ((rc = step_solve(h, sw, s))) return rc;
    if (h->fuse) {
      if ((rc = step_inverse(h, sw, s, more))) return rc;
    } else {
      if ((rc = step_inverse(h, sw, s, false))) return rc;
    }
    if (dir > 0 && sw.fl.mode == 1 && (rc = save_checkpoint(h, sw, s))) return rc;
    if (dens && (rc = density_at(h, s))) return rc;
    if (!more) break;
    if (!h->fuse && (rc = step_rhs(h, sw, s + dir))) return rc;
  }
  return WLC_OK;
}

// first segment restarted from a checkpoint (segment nstored restarts from the store)
int first_ckpt_segment(const wlc_handle* h) { return h->nstored > 0 ? h->nstored + 1 : 1; }

double* ckpt_slot(const wlc_handle* h, int j, int t) {
  return h->ckpt + ((long long)(j - first_ckpt_segment(h)) * 3 + t) * slot_size(h);
}

// restart slice jK - 2 + t of segment j >= 1
const double* restart_slice(const wlc_handle* h, int j, int t) {
  if (h->nstored > 0 && j == h->nstored)
    return h->store + (long long)(j * h->ck - 2 + t) * slot_size(h);
  return ckpt_slot(h, j, t);
}

// after forward slice s on the ring: keep the restart history of every segment boundary
// and the last slice
int save_checkpoint(wlc_handle* h, const SweepCtx& sw, int s) {
  const int K = h->ck;
  const size_t bytes = sizeof(double) * (size_t)h->Mb[block_of(h, s)] * h->hs;
  // j = floor((s + 2) / K) gives jK - 2 <= s < (j + 1)K - 2: only s <= jK is a restart slice
  const int j = (s + 2) / K;
  if (j >= first_ckpt_segment(h) && j <= h->nseg - 1 && s <= j * K)
    HIPCHK(hipMemcpyAsync(ckpt_slot(h, j, s - (j * K - 2)), fwd_slice(h, sw.fl, s), bytes,
                          hipMemcpyDeviceToDevice, sw.st));
  if (s == h->Ns)
    HIPCHK(hipMemcpyAsync(h->qlast, fwd_slice(h, sw.fl, s), bytes, hipMemcpyDeviceToDevice, sw.st));
  return WLC_OK;
}

int seg_of(const wlc_handle* h, int s) { return s == 0 ? 0 : (s - 1) / h->ck; }

FwdLoc seg_loc(const wlc_handle* h, int j) {
  FwdLoc fl;
  fl.mode = 2;
  fl.base = h->segbuf[j & 1];
  fl.lo = j == 0 ? 0 : j * h->ck - 2;
  return fl;
}

// queue the recomputation of forward segment j into buffer j % 2 on str
int recompute_segment(wlc_handle* h, int j) {
  const int b = j & 1, K = h->ck;
  if (h->seg_busy[b]) HIPCHK(hipStreamWaitEvent(h->str, h->ev_seg_free[b], 0));
  SweepCtx sw{+1, h->str, h->specr, seg_loc(h, j)};
  if (j == 0) {
    HIPCHK(wlc::launch_set_delta(h->geom, fwd_slice(h, sw.fl, 0), h->Mb[block_of(h, 0)], h->str));
  } else {
    for (int t = 0; t < 3; ++t) {
      const int s = j * K - 2 + t;
      const size_t bytes = sizeof(double) * (size_t)h->Mb[block_of(h, s)] * h->hs;
      HIPCHK(hipMemcpyAsync(fwd_slice(h, sw.fl, s), restart_slice(h, j, t), bytes,
                            hipMemcpyDeviceToDevice, h->str));
    }
  }
  int rc = sweep(h, sw, j == 0 ? 1 : j * K + 1, std::min((j + 1) * K, h->Ns), false);
  if (rc) return rc;
  HIPCHK(hipEventRecord(h->ev_seg_ready[b], h->str));
  h->seg_busy[b] = true;
  return WLC_OK;
}

// density term s (up to two blocks own slice s) into accumulator `acc` on stream `sd`
int density_launch(wlc_handle* h, int s, const double* qf, double* acc, hipStream_t sd) {
  double* rho[2] = {nullptr, nullptr};
  int lim[2] = {0, 0};
  double wt[2] = {0.0, 0.0};
  int n = 0;
  const int Ms = h->Mb[block_of(h, s)];
  for (int b = 0; b < h->p.nblk && n < 2; ++b) {
    const int j = s - h->start[b];
    if (j < 0 || j > h->p.Nblk[b]) continue;
    rho[n] = acc + (long long)b * h->hs;
    lim[n] = std::min(Ms, h->Mb[b]);
    wt[n] = h->cw[b][j];
    ++n;
  }
  if (n == 0) return WLC_OK;
  TimedLaunch t(h, 3, sd);
  HIPCHK(wlc::launch_density_acc(h->geom, qf, qs_slot(h, s), rho[0], lim[0], wt[0], rho[1],
                                 lim[1], wt[1], sd));
  return WLC_OK;
}

int density_at(wlc_handle* h, int s) {
  double* rho[2] = {nullptr, nullptr};
  int lim[2] = {0, 0};
  double wt[2] = {0.0, 0.0};
  int n = 0;
  const int Ms = h->Mb[block_of(h, s)];
  for (int b = 0; b < h->p.nblk && n < 2; ++b) {
    const int j = s - h->start[b];
    if (j < 0 || j > h->p.Nblk[b]) continue;
    rho[n] = h->rho + (long long)b * h->hs;
    lim[n] = std::min(Ms, h->Mb[b]);
    wt[n] = h->cw[b][j];
    ++n;
  }
  if (n == 0) return WLC_OK;
  const double* qf = h->q ? h->q + h->off[s] : nullptr;
  if (h->ck) {
    // forward slice s from its recomputed segment; the first density of a segment frees
    // the buffer of the segment above and queues the recomputation two segments down
    const int j = seg_of(h, s);
    if (j != h->dens_seg) {
      if (h->dens_seg >= h->nstored) {
        HIPCHK(hipEventRecord(h->ev_seg_free[h->dens_seg & 1], h->st2));
        if (j - 1 >= h->nstored) {
          int rc = recompute_segment(h, j - 1);
          if (rc) return rc;
        }
      }
      if (j >= h->nstored) HIPCHK(hipStreamWaitEvent(h->st2, h->ev_seg_ready[j & 1], 0));
      h->dens_seg = j;
    }
    qf = j < h->nstored ? h->store + (long long)s * slot_size(h) : fwd_slice(h, seg_loc(h, j), s);
  }
  HIPCHK(hipEventRecord(h->ev_step, h->st));
  HIPCHK(hipStreamWaitEvent(h->st2, h->ev_step, 0));
  {
    TimedLaunch t(h, 3, h->st2);
    HIPCHK(wlc::launch_density_acc(h->geom, qf, qs_slot(h, s), rho[0], lim[0], wt[0], rho[1],
                                   lim[1], wt[1], h->st2));
  }
  HIPCHK(hipEventRecord(h->ev_dens[s & 3], h->st2));
  h->dens_used[s & 3] = true;
  return WLC_OK;
}

// st waits for every pending density launch on st2
int join_density(wlc_handle* h) {
  for (int k = 0; k < 4; ++k)
    if (h->dens_used[k]) {
      HIPCHK(hipStreamWaitEvent(h->st, h->ev_dens[k], 0));
      h->dens_used[k] = false;
    }
  return WLC_OK;
}

void free_handle(wlc_handle* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->st) (void)hipStreamSynchronize(h->st);
  if (h->st2) (void)hipStreamSynchronize(h->st2);
  if (h->stb) (void)hipStreamSynchronize(h->stb);
  if (h->st3) (void)hipStreamSynchronize(h->st3);
  if (h->str) (void)hipStreamSynchronize(h->str);
  for (auto& r : h->recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto e : h->pool) (void)hipEventDestroy(e);
  for (auto e : h->uev) if (e) (void)hipEventDestroy(e);
  (void)hipFree(h->q);
  (void)hipFree(h->qs);
  (void)hipFree(h->spec);
  (void)hipFree(h->w);
  (void)hipFree(h->rho);
  (void)hipFree(h->zero);
  (void)hipFree(h->phi);
  (void)hipFree(h->partial);
  (void)hipFree(h->scal);
  for (auto t : h->tw) (void)hipFree(t);
  if (h->ev_step) (void)hipEventDestroy(h->ev_step);
  for (auto e : h->evf) if (e) (void)hipEventDestroy(e);
  for (auto e : h->evb) if (e) (void)hipEventDestroy(e);
  if (h->ev_pre) (void)hipEventDestroy(h->ev_pre);
  for (auto e : h->ev_join) if (e) (void)hipEventDestroy(e);
  if (h->st3) (void)hipStreamDestroy(h->st3);
  if (h->stb) (void)hipStreamDestroy(h->stb);
  (void)hipFree(h->specb);
  (void)hipFree(h->rho2);
  (void)hipFree(h->fr);
  (void)hipFree(h->store);
  (void)hipFree(h->ckpt);
  (void)hipFree(h->segbuf[0]);
  (void)hipFree(h->segbuf[1]);
  (void)hipFree(h->qlast);
  (void)hipFree(h->specr);
  for (auto e : h->ev_seg_ready) if (e) (void)hipEventDestroy(e);
  for (auto e : h->ev_seg_free) if (e) (void)hipEventDestroy(e);
  if (h->str) (void)hipStreamDestroy(h->str);
  for (auto e : h->ev_dens) if (e) (void)hipEventDestroy(e);
  if (h->st2) (void)hipStreamDestroy(h->st2);
  if (h->st) (void)hipStreamDestroy(h->st);
  delete h;
}

int validate(const wlc_params* p) {
  if (!p) return fail(WLC_EINVAL, "params is NULL");
  const int e[3] = {p->nx, p->ny, p->nz};
  for (int a = 0; a < 3; ++a) {
    if (e[a] < 2 || e[a] % 2) return fail(WLC_EINVAL, "grid edges must be even and >= 2");
    int r = e[a];
    while (r % 2 == 0) r /= 2;
    while (r % 3 == 0) r /= 3;
    if (r != 1) return fail(WLC_EINVAL, "grid edges must factor into 2s and 3s");
    if (!(p->box[a] > 0.0) || !std::isfinite(p->box[a])) return fail(WLC_EINVAL, "box edges must be positive");
  }
  if (p->nx > 256 || p->ny > 256 || p->nz > 256) return fail(WLC_EINVAL, "grid edges above 256 are not supported");
  if (!wlc::plane_supported(p->nx, p->ny))
    return fail(WLC_EINVAL, "the nx x ny half plane does not fit in LDS (ny (nx/2 + 1) must be <= 10240, or 64 x 64 / 128 x 128)");
  if (p->nblk < 1 || p->nblk > WLC_MAX_BLOCKS) return fail(WLC_EINVAL, "nblk must be 1..3");
  int tot = 0;
  for (int b = 0; b < p->nblk; ++b) {
    if (p->Nblk[b] < 1) return fail(WLC_EINVAL, "every block needs >= 1 contour step");
    if (p->Lmax[b] < 0 || p->Lmax[b] > WLC_MAX_L) return fail(WLC_EINVAL, "Lmax must be 0..8");
    if (!(p->lambda[b] >= 0.0) || !std::isfinite(p->lambda[b])) return fail(WLC_EINVAL, "lambda must be >= 0");
    tot += p->Nblk[b];
  }
  if (tot != p->Ns) return fail(WLC_EINVAL, "sum of block lengths must equal Ns");
  if (p->Ns < 3) return fail(WLC_EINVAL, "Ns must be >= 3");
  if (p->ckpt_every < 0 || (p->ckpt_every > 0 && p->ckpt_every < 3))
    return fail(WLC_EINVAL, "ckpt_every must be 0 (automatic) or >= 3");
  if (p->spectrum_fp32 != 0 && p->spectrum_fp32 != 1)
    return fail(WLC_EINVAL, "spectrum_fp32 must be 0 or 1");
  for (int r : p->reserved)
    if (r) return fail(WLC_EINVAL, "reserved fields must be zero");
  return WLC_OK;
}

int pick_tx(int nz) {
  // generic kernel B: one thread per k-point, TX kx columns of nz points per workgroup;
  // at most four columns (measured: 24^3 and 32^3 at Lmax 4, 48^3 at Lmax 6 run fastest
  // with four) and at least one wave per workgroup
  if (nz >= 256) return 1;
  const int tx = std::min(4, 256 / nz);
  return std::max(tx, (64 + nz - 1) / nz);
}

// backward sweep concurrent with the queued forward sweep (see the handle notes)
int backward_concurrent(wlc_handle* h) {
  const int Ns = h->Ns;
  const bool dens = h->fwd_done;
  if (!h->pre_recorded) HIPCHK(hipEventRecord(h->ev_pre, h->st));
  h->pre_recorded = false;
  HIPCHK(hipStreamWaitEvent(h->stb, h->ev_pre, 0));
  const long long nacc = (long long)h->p.nblk * h->hs;
  HIPCHK(wlc::launch_fill(h->rho, 0.0, nacc, h->stb));
  HIPCHK(wlc::launch_fill(h->rho2, 0.0, nacc, h->stb));
  HIPCHK(wlc::launch_set_delta(h->geom, qs_slot(h, Ns), h->Mb[block_of(h, Ns)], h->stb));
  HIPCHK(hipEventRecord(h->evb[Ns], h->stb));
  int rc = sweep(h, main_sweep(h, -1), Ns - 1, 0, false);
  if (rc == WLC_OK && dens) {
    const int mid = Ns / 2;
    // s <= Ns/2: ready when the backward sweep reaches s, in decreasing s
    for (int s = mid; s >= 0 && rc == WLC_OK; --s) {
      HIPCHK(hipStreamWaitEvent(h->st2, h->evf[s], 0));
      HIPCHK(hipStreamWaitEvent(h->st2, h->evb[s], 0));
      rc = density_launch(h, s, h->q + h->off[s], h->rho, h->st2);
    }
    // s > Ns/2: ready when the forward sweep reaches s, in increasing s
    for (int s = mid + 1; s <= Ns && rc == WLC_OK; ++s) {
      HIPCHK(hipStreamWaitEvent(h->st3, h->evf[s], 0));
      HIPCHK(hipStreamWaitEvent(h->st3, h->evb[s], 0));
      rc = density_launch(h, s, h->q + h->off[s], h->rho2, h->st3);
    }
  }
  // join every stream into st, also on failure
  const hipStream_t side[3] = {h->stb, h->st2, h->st3};
  for (int k = 0; k < 3; ++k) {
    HIPCHK(hipEventRecord(h->ev_join[k], side[k]));
    HIPCHK(hipStreamWaitEvent(h->st, h->ev_join[k], 0));
  }
  if (rc) return rc;
  if (dens) HIPCHK(wlc::launch_axpy(h->rho, 1.0, h->rho2, nacc, h->st));
  h->bwd_done = dens;
  return WLC_OK;
}

// ---------------------------------------------------------------------------------------
// One contour step on caller device buffers in the dense layout (z-plane stride nx ny, field
// stride nz nx ny): kernels A, B and C of the sweep on the target s of direction dir, with
// the chain's block rules (target block's field, Lmax and lambda; history zero-padded or
// truncated to the target block's harmonics).  `spec` (half spectrum of the target's M
// harmonics) and `zero` (one zero field) are the scratch.

int check_step_args(const wlc_params* p, int dir, int s, const double* const hist[3],
                    const double* w, const double* out, int* order) {
  if (dir != 1 && dir != -1) return fail(WLC_EINVAL, "dir must be +1 or -1");
  if (!hist || !w || !out) return fail(WLC_EINVAL, "NULL argument");
  if ((dir > 0 && (s < 1 || s > p->Ns)) || (dir < 0 && (s < 0 || s > p->Ns - 1)))
    return fail(WLC_EINVAL, "target s out of range for this direction");
  *order = dir > 0 ? std::min(s, 3) : std::min(p->Ns - s, 3);
  for (int j = 0; j < *order; ++j)
    if (!hist[j]) return fail(WLC_EINVAL, "a history slice within the BDF order is NULL");
  return WLC_OK;
}

int params_block_of(const wlc_params& p, int s) {
  int st = 0;
  for (int b = 0; b < p.nblk; ++b) {
    st += p.Nblk[b];
    if (s <= st) return b;
  }
  return p.nblk - 1;
}

int queue_dense_step(const wlc_params& p, const wlc::Geom& g, int dir, int s,
                     int order, const double* const hist[3], const double* w, double* out,
                     double2* spec, const double* zero, hipStream_t st) {
  const double ds = 1.0 / p.Ns;
  const int blk = params_block_of(p, s);
  const int M = (p.Lmax[blk] + 1) * (p.Lmax[blk] + 1);
  wlc::RhsArgs ra{};
  double cdelta;
  bdf_coeffs(order, ra.a, ra.b, &cdelta);
  for (int j = 0; j < 3; ++j) {
    const int sh = s - dir * (j + 1);
    if (j < order) {
      const int hb = params_block_of(p, sh);
      ra.h[j] = hist[j];
      ra.hM[j] = (p.Lmax[hb] + 1) * (p.Lmax[hb] + 1);
    } else {
      ra.h[j] = nullptr;
      ra.hM[j] = 0;
    }
  }
  ra.ds = ds;
  ra.w = w + (long long)blk * g.hs;
  ra.zero = zero;
  ra.M = M;
  ra.spec = spec;
  HIPCHK(wlc::launch_rhs_fft(g, ra, order, st));
  wlc::SolveArgs sa{};
  sa.spec = spec;
  sa.M = M;
  sa.L = p.Lmax[blk];
  const wlc::ZOpts o = wlc::zsolve_opts(p.nz, p.Lmax[blk], p.spectrum_fp32);
  sa.TX = o.tx;
  sa.cdelta = cdelta;
  sa.sigma = dir > 0 ? 1.0 : -1.0;
  sa.ds = ds;
  sa.dslam = ds * p.lambda[blk];
  for (int a = 0; a < 3; ++a) sa.kfac[a] = 2.0 * M_PI / p.box[a];
  sa.solve = 1;
  sa.generic = o.generic;
  sa.occ = o.occ;
  sa.split = o.split;
  HIPCHK(wlc::launch_zsolve(g, sa, st));
  wlc::InvArgs ia{};
  ia.spec = spec;
  ia.M = M;
  ia.scale = 1.0 / ((double)p.nx * p.ny * p.nz);
  ia.out = out;
  HIPCHK(wlc::launch_ifft_store(g, ia, st));
  return WLC_OK;
}

// orientation tables built on the host and checked against the compiled-in ones, once per
// process and once per device
int ensure_tables(int device) {
  std::lock_guard<std::mutex> lk(g_tab_mu);
  if (!g_tab_built) {
    std::string err = wlc::build_orientation_tables(wlc::kMaxL, &g_tab);
    if (!err.empty()) return fail(WLC_EINVAL, "orientation tables: " + err);
    g_tab_built = true;
  }
  if ((int)g_tab_uploaded.size() <= device) g_tab_uploaded.resize(device + 1, 0);
  if (!g_tab_uploaded[device]) {
    if (wlc::upload_orientation_tables(g_tab) != hipSuccess)
      return fail(WLC_EHIP, "compiled-in orientation tables differ from the host set-up (regenerate wlc_tables.inc)");
    g_tab_uploaded[device] = 1;
  }
  return WLC_OK;
}

// ---------------------------------------------------------------------------------------
// Scratch of wlc_mde_step_noalloc: a fixed device array, one copy per GPU, held until the
// process ends.  Layout: FFT twiddles of the three axes at offset 0 (3 x 256 complex), one
// zero field at kScrZero, the half spectrum at kScrZero + the zero field rounded up to
// 256 bytes.  The host reaches each device's copy only through the address
// hipGetSymbolAddress returns on that device.
constexpr size_t kScratchBytes = size_t(128) << 20;
constexpr size_t kScrZero = 64 << 10;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scr_mu;
std::vector<unsigned char*> g_scr_ptr;  // per device, nullptr until resolved

int scratch_on(int device, unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_scr_mu);
  if ((int)g_scr_ptr.size() <= device) g_scr_ptr.resize(device + 1, nullptr);
  if (!g_scr_ptr[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(WLC_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scr_ptr[device] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scr_ptr[device];
  return WLC_OK;
}

// host twiddle tables per edge, kept for the life of the process (the source of the
// per-call upload stays valid however the runtime stages the copy)
std::mutex g_tw_mu;
std::vector<std::vector<double2>> g_tw_host(257);

const double2* host_twiddles(const wlc::AxisPlan& pl) {
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto& v = g_tw_host[pl.n];
  if (v.empty()) {
    v.resize(pl.n);
    for (int t = 0; t < pl.n; ++t) v[t] = make_double2(pl.tw[2 * t], pl.tw[2 * t + 1]);
  }
  return v.data();
}

}  // namespace

namespace wlc {
int host_fail(int code, const std::string& msg) { return fail(code, msg); }
int validate_params(const wlc_params* p) { return validate(p); }
void simpson_weights(int n, double ds, std::vector<double>* w) { contour_weights(n, ds, w); }
int zsolve_tx(int nz) { return pick_tx(nz); }
int orientation_tables_ready(int device) { return ensure_tables(device); }
void bdf(int order, double a[3], double b[3], double* cdelta) { bdf_coeffs(order, a, b, cdelta); }
}  // namespace wlc

extern "C" {

const char* wlc_mde_last_error(void) { return g_last_error.c_str(); }

int wlc_mde_create(const wlc_params* p, wlc_handle** out) {
  if (!out) return fail(WLC_EINVAL, "out is NULL");
  *out = nullptr;
  int rc = validate(p);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(g_tab_mu);
    if (!g_tab_built) {
      std::string err = wlc::build_orientation_tables(wlc::kMaxL, &g_tab);
      if (!err.empty()) return fail(WLC_EINVAL, "orientation tables: " + err);
      g_tab_built = true;
    }
  }
  wlc_handle* h = new wlc_handle();
  h->p = *p;
  if (const char* fu = std::getenv("WLC_FUSE")) h->fuse = std::atoi(fu) != 0;
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    return fail(WLC_EHIP, "no HIP device");
  }
  if ((rc = ensure_tables(h->device))) {
    delete h;
    return rc;
  }
  h->N = (long long)p->nx * p->ny * p->nz;
  // padded real-space strides (see layout notes in wlc_kernels.h)
  h->zs = (long long)p->nx * p->ny + 16;
  h->hs = (long long)p->nz * h->zs + 32;
  h->nxh = p->nx / 2 + 1;
  h->Nk = (long long)h->nxh * p->ny * p->nz;
  h->Ns = p->Ns;
  h->ds = 1.0 / p->Ns;  // initialize_mod.f90:140
  int st = 0;
  for (int b = 0; b < p->nblk; ++b) {
    h->Lb[b] = p->Lmax[b];
    h->Mb[b] = (p->Lmax[b] + 1) * (p->Lmax[b] + 1);
    h->start[b] = st;
    st += p->Nblk[b];
    h->Mmax = std::max(h->Mmax, h->Mb[b]);
    std::vector<double> cw;
    contour_weights(p->Nblk[b], h->ds, &cw);
    h->cw.push_back(cw);
  }
  h->off.resize(h->Ns + 2);
  h->off[0] = 0;
  for (int s = 0; s <= h->Ns; ++s) h->off[s + 1] = h->off[s] + (long long)h->Mb[block_of(h, s)] * h->hs;
  const int edges[3] = {p->nx, p->ny, p->nz};
  for (int a = 0; a < 3; ++a) {
    std::string err = wlc::make_axis_plan(edges[a], &h->plan[a]);
    if (!err.empty()) { free_handle(h); return fail(WLC_EINVAL, err); }
    if (h->plan[a].npass > 8) { free_handle(h); return fail(WLC_EINVAL, "too many FFT passes"); }
  }
#define ALLOC(ptr, bytes)                                                        \
  do {                                                                           \
    if (hipMalloc((void**)&(ptr), (bytes)) != hipSuccess) {                      \
      free_handle(h);                                                            \
      return fail(WLC_ENOMEM, std::string("device allocation failed: ") + #ptr); \
    }                                                                            \
  } while (0)
  if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&h->st2, hipStreamNonBlocking) != hipSuccess) {
    free_handle(h);
    return fail(WLC_EHIP, "stream creation failed");
  }
  bool ev_ok = hipEventCreateWithFlags(&h->ev_step, hipEventDisableTiming) == hipSuccess;
  for (auto& e : h->ev_dens) ev_ok = ev_ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
  if (!ev_ok) {
    free_handle(h);
    return fail(WLC_EHIP, "event creation failed");
  }
  const size_t store_bytes = sizeof(double) * (size_t)h->off[h->Ns + 1];
  // sized for FP64 elements in either mode, so that wlc_mde_set_spectrum_fp32 can switch
  const size_t spec_bytes = sizeof(double2) * (size_t)h->Mmax * p->ny * wlc::kx_tiles(h->nxh) * wlc::kKxTile * p->nz;
  const size_t slot_bytes = sizeof(double) * (size_t)slot_size(h);
  const size_t field_bytes = sizeof(double) * (size_t)p->nblk * h->hs;
  {
    // full forward store when it fits (with the backward ring, spectrum, fields and
    // 2 GiB to spare); otherwise segments of ckpt_every (default 16) contour steps
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    const size_t base = 4 * slot_bytes + spec_bytes + 4 * field_bytes + (size_t(2) << 30);
    bool automatic = false;
    if (p->ckpt_every > 0) h->ck = p->ckpt_every;
    else if (store_bytes + base > free_b) h->ck = 16, automatic = true;
    if (h->ck) {
      h->nseg = (h->Ns + h->ck - 1) / h->ck;
      // automatic mode: keep as many of the lowest segments whole as fit beside the
      // checkpoint buffers (the last segment is always recomputed); an explicit ckpt_every
      // checkpoints every segment unless WLC_STORED_SEGMENTS asks for a mix (tests)
      const size_t ck_fixed = base + (size_t)(4 + 1 + 2 * (h->ck + 3)) * slot_bytes + spec_bytes;
      if (automatic) {
        for (int n = h->nseg - 1; n >= 1; --n) {
          const size_t need = ck_fixed + (size_t)(n * h->ck + 1) * slot_bytes +
                              (size_t)(h->nseg - (n + 1)) * 3 * slot_bytes;
          if (need <= free_b) {
            h->nstored = n;
            break;
          }
        }
      } else if (const char* ns = std::getenv("WLC_STORED_SEGMENTS")) {
        h->nstored = std::max(0, std::min(std::atoi(ns), h->nseg - 1));
      }
    }
  }
  if (h->ck) {
    ALLOC(h->fr, 4 * slot_bytes);
    if (h->nstored > 0) ALLOC(h->store, (size_t)(h->nstored * h->ck + 1) * slot_bytes);
    {
      const int nck = h->nseg - (h->nstored > 0 ? h->nstored + 1 : 1);
      if (nck > 0) ALLOC(h->ckpt, (size_t)nck * 3 * slot_bytes);
    }
    for (int b = 0; b < 2; ++b) ALLOC(h->segbuf[b], (size_t)(h->ck + 3) * slot_bytes);
    ALLOC(h->qlast, slot_bytes);
    ALLOC(h->specr, spec_bytes);
    if (hipStreamCreateWithFlags(&h->str, hipStreamNonBlocking) != hipSuccess) {
      free_handle(h);
      return fail(WLC_EHIP, "stream creation failed");
    }
    bool ok = true;
    for (auto& e : h->ev_seg_ready) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    for (auto& e : h->ev_seg_free) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
      free_handle(h);
      return fail(WLC_EHIP, "event creation failed");
    }
  } else {
    ALLOC(h->q, store_bytes);
  }
  if (!h->ck) {
    // concurrent sweeps need the full backward store, a second spectrum and a second
    // accumulator; take them when they fit with 2 GiB to spare
    size_t free_b = 0, total_b = 0;
    const size_t extra = (p->keep_qstar ? 0 : store_bytes) + spec_bytes +
                         sizeof(double) * (size_t)p->nblk * h->hs;
    const char* ce = std::getenv("WLC_CONCURRENT");
    const bool want = !(ce && std::atoi(ce) == 0);
    h->conc = want && hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
              free_b > extra + spec_bytes + (size_t(2) << 30);
  }
  if (p->keep_qstar || h->conc) ALLOC(h->qs, store_bytes);
  else ALLOC(h->qs, sizeof(double) * (size_t)4 * h->Mmax * h->hs);
  ALLOC(h->spec, spec_bytes);
  if (h->conc) {
    ALLOC(h->specb, spec_bytes);
    ALLOC(h->rho2, sizeof(double) * (size_t)p->nblk * h->hs);
    if (hipStreamCreateWithFlags(&h->stb, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&h->st3, hipStreamNonBlocking) != hipSuccess) {
      free_handle(h);
      return fail(WLC_EHIP, "stream creation failed");
    }
    h->evf.assign(h->Ns + 1, nullptr);
    h->evb.assign(h->Ns + 1, nullptr);
    bool ok = hipEventCreateWithFlags(&h->ev_pre, hipEventDisableTiming) == hipSuccess;
    for (auto& e : h->ev_join) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    for (auto& e : h->evf) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    for (auto& e : h->evb) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
      free_handle(h);
      return fail(WLC_EHIP, "event creation failed");
    }
  }
  ALLOC(h->w, sizeof(double) * (size_t)p->nblk * h->hs);
  ALLOC(h->zero, sizeof(double) * (size_t)h->hs);
  if (hipMemset(h->zero, 0, sizeof(double) * (size_t)h->hs) != hipSuccess) {
    free_handle(h);
    return fail(WLC_EHIP, "zero field initialisation failed");
  }
  ALLOC(h->rho, sizeof(double) * (size_t)p->nblk * h->hs);
  ALLOC(h->phi, sizeof(double) * (size_t)p->nblk * h->hs);
  ALLOC(h->partial, sizeof(double) * h->nparts);
  ALLOC(h->scal, sizeof(double) * 4);
  for (int a = 0; a < 3; ++a) {
    ALLOC(h->tw[a], sizeof(double2) * (size_t)edges[a]);
    if (hipMemcpy(h->tw[a], h->plan[a].tw.data(), sizeof(double2) * edges[a], hipMemcpyHostToDevice) != hipSuccess) {
      free_handle(h);
      return fail(WLC_EHIP, "twiddle upload failed");
    }
  }
#undef ALLOC
  wlc::Geom& g = h->geom;
  g.nx = p->nx; g.ny = p->ny; g.nz = p->nz; g.nxh = h->nxh;
  g.nt = wlc::kx_tiles(h->nxh);
  g.spec32 = p->spectrum_fp32;
  g.zs = h->zs;
  g.hs = h->hs;
  wlc::AxisDev* ax[3] = {&g.ax, &g.ay, &g.az};
  for (int a = 0; a < 3; ++a) {
    ax[a]->n = h->plan[a].n;
    ax[a]->npass = h->plan[a].npass;
    for (int i = 0; i < 8; ++i) ax[a]->radix[i] = i < h->plan[a].npass ? h->plan[a].radix[i] : 0;
    ax[a]->tw = h->tw[a];
  }
  *out = h;
  g_last_error.clear();
  return WLC_OK;
}

void wlc_mde_destroy(wlc_handle* h) { free_handle(h); }

int wlc_mde_set_box(wlc_handle* h, const double box[3]) {
  if (!h || !box) return fail(WLC_EINVAL, "NULL argument");
  for (int a = 0; a < 3; ++a)
    if (!(box[a] > 0.0) || !std::isfinite(box[a])) return fail(WLC_EINVAL, "box edges must be positive");
  for (int a = 0; a < 3; ++a) h->p.box[a] = box[a];
  h->fwd_done = h->bwd_done = false;
  h->pre_recorded = false;
  return WLC_OK;
}

int wlc_mde_set_w(wlc_handle* h, const double* w_host) {
  if (!h || !w_host) return fail(WLC_EINVAL, "NULL argument");
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(copy_fields_in(h, h->w, w_host, h->p.nblk, hipMemcpyHostToDevice));
  HIPCHK(hipStreamSynchronize(h->st));
  h->have_w = true;
  h->fwd_done = h->bwd_done = false;
  h->pre_recorded = false;  // a concurrent backward sweep must start after the copy
  return WLC_OK;
}

int wlc_mde_set_w_device(wlc_handle* h, const double* w_dev) {
  if (!h || !w_dev) return fail(WLC_EINVAL, "NULL argument");
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(copy_fields_in(h, h->w, w_dev, h->p.nblk, hipMemcpyDeviceToDevice));
  h->have_w = true;
  h->fwd_done = h->bwd_done = false;
  h->pre_recorded = false;  // a concurrent backward sweep must start after the copy
  return WLC_OK;
}

int wlc_mde_forward(wlc_handle* h) {
  if (!h) return fail(WLC_EINVAL, "NULL handle");
  if (!h->have_w) return fail(WLC_ESTATE, "set_w must precede forward");
  HIPCHK(hipSetDevice(h->device));
  // bound the pending timing records of callers that never read them back
  if (h->timing && h->recs.size() > 16384) {
    const int rt = flush_timing(h);
    if (rt) return rt;
  }
  if (h->conc) {
    HIPCHK(hipEventRecord(h->ev_pre, h->st));  // the backward sweep may start from here
    h->pre_recorded = true;
  }
  const SweepCtx sw = main_sweep(h, +1);
  HIPCHK(wlc::launch_set_delta(h->geom, fwd_slice(h, sw.fl, 0), h->Mb[block_of(h, 0)], h->st));
  if (h->conc) HIPCHK(hipEventRecord(h->evf[0], h->st));
  int rc = sweep(h, sw, 1, h->Ns, false);
  if (rc) return rc;
  h->fwd_done = true;
  h->bwd_done = false;
  // timing mode 2: finish the forward sweep before anything else is queued, so that each
  // step kernel is timed with the GPU to itself
  if (h->timing_serial) return flush_timing(h);
  return WLC_OK;
}

int wlc_mde_backward(wlc_handle* h) {
  if (!h) return fail(WLC_EINVAL, "NULL handle");
  if (!h->have_w) return fail(WLC_ESTATE, "set_w must precede backward");
  HIPCHK(hipSetDevice(h->device));
  const int Ns = h->Ns;
  if (h->conc) return backward_concurrent(h);
  HIPCHK(wlc::launch_fill(h->rho, 0.0, (long long)h->p.nblk * h->hs, h->st));
  HIPCHK(wlc::launch_set_delta(h->geom, qs_slot(h, Ns), h->Mb[block_of(h, Ns)], h->st));
  const bool dens = h->fwd_done;
  int rc = WLC_OK;
  if (dens && h->ck) {
    // the recomputation stream starts after everything queued so far on st
    HIPCHK(hipEventRecord(h->ev_step, h->st));
    HIPCHK(hipStreamWaitEvent(h->str, h->ev_step, 0));
    h->dens_seg = -1;
    h->seg_busy[0] = h->seg_busy[1] = false;
    const int J = h->nseg - 1;  // >= nstored: the store never holds the last slice
    rc = recompute_segment(h, J);
    if (rc == WLC_OK && J - 1 >= h->nstored) rc = recompute_segment(h, J - 1);
  }
  if (rc == WLC_OK && dens) rc = density_at(h, Ns);
  if (rc == WLC_OK) rc = sweep(h, main_sweep(h, -1), Ns - 1, 0, dens);
  const int rj = join_density(h);  // also on failure: no density launch outlives the call
  if (h->ck) {
    HIPCHK(hipEventRecord(h->ev_seg_ready[0], h->str));
    HIPCHK(hipStreamWaitEvent(h->st, h->ev_seg_ready[0], 0));
  }
  if (rc) return rc;
  if (rj) return rj;
  h->bwd_done = dens;
  return WLC_OK;
}

static int density_common(wlc_handle* h, double* Q, double* totden) {
  if (!h->fwd_done || !h->bwd_done) return fail(WLC_ESTATE, "density needs a forward and a backward sweep on the current w");
  HIPCHK(hipSetDevice(h->device));
  const double dV = (h->p.box[0] / h->p.nx) * (h->p.box[1] / h->p.ny) * (h->p.box[2] / h->p.nz);
  const double V = h->p.box[0] * h->p.box[1] * h->p.box[2];
  HIPCHK(wlc::launch_weighted_sum(h->geom, fwd_last(h), 1, dV, h->partial, h->nparts, h->scal, h->st));
  HIPCHK(wlc::launch_weighted_sum(h->geom, h->rho, h->p.nblk, dV, h->partial, h->nparts, h->scal + 1, h->st));
  HIPCHK(wlc::launch_phi(h->geom, h->rho, h->p.nblk, V, h->scal + 1, h->phi, h->st));
  double sc[2];
  HIPCHK(hipMemcpyAsync(sc, h->scal, sizeof(sc), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  // timed launches are read out here, where the host waits anyway, so that timing does not
  // add a synchronisation between the forward and the backward sweep
  if (h->timing) {
    const int rt = flush_timing(h);
    if (rt) return rt;
  }
  const double Qv = sc[0] / V;
  if (!std::isfinite(Qv) || !std::isfinite(sc[1]) || sc[1] == 0.0) return fail(WLC_ENONFINITE, "Q or totden not finite");
  if (Q) *Q = Qv;
  if (totden) *totden = sc[1];
  return WLC_OK;
}

int wlc_mde_density(wlc_handle* h, double* phi_host, double* Q, double* totden) {
  if (!h) return fail(WLC_EINVAL, "NULL handle");
  int rc = density_common(h, Q, totden);
  if (rc) return rc;
  if (phi_host) {
    HIPCHK(copy_fields_out(h, phi_host, h->phi, h->p.nblk, hipMemcpyDeviceToHost));
    HIPCHK(hipStreamSynchronize(h->st));
    for (long long e = 0; e < (long long)h->p.nblk * h->N; ++e)
      if (!std::isfinite(phi_host[e])) return fail(WLC_ENONFINITE, "phi not finite");
  }
  return WLC_OK;
}

int wlc_mde_density_device(wlc_handle* h, double* phi_dev, double* Q, double* totden) {
  if (!h) return fail(WLC_EINVAL, "NULL handle");
  int rc = density_common(h, Q, totden);
  if (rc) return rc;
  if (phi_dev) HIPCHK(copy_fields_out(h, phi_dev, h->phi, h->p.nblk, hipMemcpyDeviceToDevice));
  return WLC_OK;
}

int wlc_mde_get_q(wlc_handle* h, int dir, int s, double* q_host) {
  if (!h || !q_host) return fail(WLC_EINVAL, "NULL argument");
  if (s < 0 || s > h->Ns || (dir != 1 && dir != -1)) return fail(WLC_EINVAL, "bad slice");
  HIPCHK(hipSetDevice(h->device));
  const long long M = h->Mb[block_of(h, s)];
  const double* src;
  if (dir > 0) {
    if (!h->fwd_done) return fail(WLC_ESTATE, "no forward sweep");
    if (h->ck && s != h->Ns && (h->nstored == 0 || s > h->nstored * h->ck))
      return fail(WLC_ESTATE, "this forward slice is not stored in checkpoint mode");
    src = !h->ck ? h->q + h->off[s] : s == h->Ns ? h->qlast : h->store + (long long)s * slot_size(h);
  } else {
    if (!h->p.keep_qstar && s != 0) return fail(WLC_ESTATE, "backward slices other than 0 need keep_qstar");
    src = qs_slot(h, s);
  }
  HIPCHK(hipStreamSynchronize(h->st));
  HIPCHK(copy_fields_out(h, q_host, src, (int)M, hipMemcpyDeviceToHost));
  HIPCHK(hipStreamSynchronize(h->st));
  return WLC_OK;
}

int wlc_mde_slice_M(const wlc_handle* h, int s) {
  if (!h || s < 0 || s > h->Ns) return WLC_EINVAL;
  return h->Mb[block_of(h, s)];
}

int wlc_mde_forward_stored(const wlc_handle* h) {
  if (!h) return fail(WLC_EINVAL, "NULL handle");
  if (!h->ck) return h->Ns + 1;
  return h->nstored > 0 ? h->nstored * h->ck + 1 : 0;
}

int wlc_mde_block_of(const wlc_handle* h, int s) {
  if (!h || s < 0 || s > h->Ns) return WLC_EINVAL;
  return block_of(h, s);
}

int wlc_mde_sync(wlc_handle* h) {
  if (!h) return fail(WLC_EINVAL, "NULL handle");
  HIPCHK(hipStreamSynchronize(h->st));
  return WLC_OK;
}

void* wlc_mde_stream(wlc_handle* h) { return h ? (void*)h->st : nullptr; }

int wlc_mde_set_spectrum_fp32(wlc_handle* h, int on) {
  if (!h) return fail(WLC_EINVAL, "NULL handle");
  if (on != 0 && on != 1) return fail(WLC_EINVAL, "spectrum_fp32 must be 0 or 1");
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamSynchronize(h->st));  // no queued sweep still uses the old element type
  h->p.spectrum_fp32 = on;
  h->geom.spec32 = on;
  h->fwd_done = h->bwd_done = false;  // earlier sweeps are not reused across the switch
  return WLC_OK;
}

int wlc_mde_set_timing(wlc_handle* h, int enable) {
  if (!h) return fail(WLC_EINVAL, "NULL handle");
  int rc = flush_timing(h);
  if (rc) return rc;
  if (enable < 0 || enable > 2) return fail(WLC_EINVAL, "timing mode must be 0, 1 or 2");
  h->timing = enable != 0;
  h->timing_serial = enable == 2;
  for (int k = 0; k < 5; ++k) { h->ms[k] = 0.0; h->cnt[k] = 0; }
  return WLC_OK;
}

int wlc_mde_get_timing(wlc_handle* h, double* ms, long long* count, int nk) {
  if (!h) return fail(WLC_EINVAL, "NULL handle");
  int rc = flush_timing(h);
  if (rc) return rc;
  for (int k = 0; k < nk && k < 5; ++k) {
    if (ms) ms[k] = h->ms[k];
    if (count) count[k] = h->cnt[k];
  }
  return WLC_OK;
}

int wlc_mde_event_record(wlc_handle* h, int slot) {
  if (!h || slot < 0 || slot >= 8) return fail(WLC_EINVAL, "bad event slot");
  if (!h->uev[slot]) HIPCHK(hipEventCreate(&h->uev[slot]));
  HIPCHK(hipEventRecord(h->uev[slot], h->st));
  return WLC_OK;
}

int wlc_mde_event_elapsed(wlc_handle* h, int a, int b, double* ms) {
  if (!h || a < 0 || a >= 8 || b < 0 || b >= 8 || !h->uev[a] || !h->uev[b] || !ms)
    return fail(WLC_EINVAL, "bad event slots");
  HIPCHK(hipEventSynchronize(h->uev[b]));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, h->uev[a], h->uev[b]));
  *ms = t;
  return WLC_OK;
}

int wlc_test_solve(int L, int nk, const double* kvec_host, double cdelta, double sigma, double ds,
                   double dslam, double* v_host) {
  if (L < 0 || L > WLC_MAX_L || nk <= 0 || !kvec_host || !v_host) return fail(WLC_EINVAL, "bad arguments");
  {
    std::lock_guard<std::mutex> lk(g_tab_mu);
    if (!g_tab_built) {
      std::string err = wlc::build_orientation_tables(wlc::kMaxL, &g_tab);
      if (!err.empty()) return fail(WLC_EINVAL, "orientation tables: " + err);
      g_tab_built = true;
    }
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if ((int)g_tab_uploaded.size() <= dev) g_tab_uploaded.resize(dev + 1, 0);
    if (!g_tab_uploaded[dev]) {
      HIPCHK(wlc::upload_orientation_tables(g_tab));
      g_tab_uploaded[dev] = 1;
    }
  }
  const int M = (L + 1) * (L + 1);
  double* dk = nullptr;
  double2* dv = nullptr;
  HIPCHK(hipMalloc((void**)&dk, sizeof(double) * 3 * (size_t)nk));
  if (hipMalloc((void**)&dv, sizeof(double2) * (size_t)nk * M) != hipSuccess) {
    (void)hipFree(dk);
    return fail(WLC_ENOMEM, "test buffer allocation failed");
  }
  int rc = WLC_OK;
  hipError_t e = hipMemcpy(dk, kvec_host, sizeof(double) * 3 * (size_t)nk, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dv, v_host, sizeof(double2) * (size_t)nk * M, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = wlc::launch_test_solve(L, nk, dk, cdelta, sigma, ds, dslam, dv, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(v_host, dv, sizeof(double2) * (size_t)nk * M, hipMemcpyDeviceToHost);
  if (e != hipSuccess) rc = fail(WLC_EHIP, std::string("test solve: ") + hipGetErrorString(e));
  (void)hipFree(dk);
  (void)hipFree(dv);
  return rc;
}

int wlc_mde_step(wlc_handle* h, int dir, int s, const double* const hist_dev[3],
                 const double* w_dev, double* out_dev) {
  if (!h) return fail(WLC_EINVAL, "NULL handle");
  int order = 0;
  int rc = check_step_args(&h->p, dir, s, hist_dev, w_dev, out_dev, &order);
  if (rc) return rc;
  HIPCHK(hipSetDevice(h->device));
  wlc::Geom g = h->geom;
  g.zs = (long long)h->p.nx * h->p.ny;
  g.hs = (long long)h->p.nz * g.zs;
  return queue_dense_step(h->p, g, dir, s, order, hist_dev, w_dev, out_dev, h->spec, h->zero, h->st);
}

int wlc_mde_step_noalloc(const wlc_params* p, int dir, int s, const double* const hist_dev[3],
                         const double* w_dev, double* out_dev, void* stream) {
  int rc = validate(p);
  if (rc) return rc;
  int order = 0;
  if ((rc = check_step_args(p, dir, s, hist_dev, w_dev, out_dev, &order))) return rc;
  const int nxh = p->nx / 2 + 1;
  const int blk = params_block_of(*p, s);
  const size_t M = (size_t)(p->Lmax[blk] + 1) * (p->Lmax[blk] + 1);
  const size_t field = (size_t)p->nx * p->ny * p->nz;
  const size_t zero_bytes = (sizeof(double) * field + 255) / 256 * 256;
  const size_t spec_bytes = sizeof(double2) * M * p->ny * wlc::kx_tiles(nxh) * wlc::kKxTile * p->nz;
  if (kScrZero + zero_bytes + spec_bytes > kScratchBytes)
    return fail(WLC_EINVAL, "step too large for the fixed scratch (" + std::to_string(kScratchBytes >> 20) +
                                " MiB: one zero field plus the half spectrum of the target's harmonics)");
  wlc::AxisPlan plan[3];
  const int edges[3] = {p->nx, p->ny, p->nz};
  for (int a = 0; a < 3; ++a) {
    std::string err = wlc::make_axis_plan(edges[a], &plan[a]);
    if (!err.empty()) return fail(WLC_EINVAL, err);
    if (plan[a].npass > 8) return fail(WLC_EINVAL, "too many FFT passes");
  }
  int device = 0;
  HIPCHK(hipGetDevice(&device));
  if ((rc = ensure_tables(device))) return rc;
  unsigned char* scr = nullptr;
  if ((rc = scratch_on(device, &scr))) return rc;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  double2* tw = reinterpret_cast<double2*>(scr);
  double* zero = reinterpret_cast<double*>(scr + kScrZero);
  double2* spec = reinterpret_cast<double2*>(scr + kScrZero + zero_bytes);
  wlc::Geom g{};
  g.nx = p->nx; g.ny = p->ny; g.nz = p->nz; g.nxh = nxh;
  g.nt = wlc::kx_tiles(nxh);
  g.spec32 = p->spectrum_fp32;
  g.zs = (long long)p->nx * p->ny;
  g.hs = (long long)p->nz * g.zs;
  wlc::AxisDev* ax[3] = {&g.ax, &g.ay, &g.az};
  for (int a = 0; a < 3; ++a) {
    ax[a]->n = plan[a].n;
    ax[a]->npass = plan[a].npass;
    for (int i = 0; i < 8; ++i) ax[a]->radix[i] = i < plan[a].npass ? plan[a].radix[i] : 0;
    ax[a]->tw = tw + 256 * a;
    HIPCHK(hipMemcpyAsync(tw + 256 * a, host_twiddles(plan[a]), sizeof(double2) * plan[a].n,
                          hipMemcpyHostToDevice, st));
  }
  // the array keeps the previous call's contents: the zero field is cleared every call
  HIPCHK(wlc::launch_fill(zero, 0.0, (long long)field, st));
  return queue_dense_step(*p, g, dir, s, order, hist_dev, w_dev, out_dev, spec, zero, st);
}

}  // extern "C"
