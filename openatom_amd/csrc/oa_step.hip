// oa_step.hip -- context, layout and the C ABI of the CP electronic-step
// library.  Host code only; the kernels live in oa_fft.hip and oa_gemm.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "oa_ctx_info.h"
#include "oa_internal.h"
#include "oa_kernels.h"
#include "oa_step.h"

using oa::FftCtx;
using oa::StepDev;

namespace {

constexpr double kTimeConv = 0.0241888;  // piny_constants.h:58 (fs -> a.u.)
constexpr double kCpEmagic = 0.00050;    // piny_constants.h:59

double inv3(const double h[9], double hi[9]) {
  const double det = h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
                     h[2] * (h[3] * h[7] - h[4] * h[6]);
  const double id = 1.0 / det;
  hi[0] = (h[4] * h[8] - h[5] * h[7]) * id;
  hi[1] = (h[2] * h[7] - h[1] * h[8]) * id;
  hi[2] = (h[1] * h[5] - h[2] * h[4]) * id;
  hi[3] = (h[5] * h[6] - h[3] * h[8]) * id;
  hi[4] = (h[0] * h[8] - h[2] * h[6]) * id;
  hi[5] = (h[2] * h[3] - h[0] * h[5]) * id;
  hi[6] = (h[3] * h[7] - h[4] * h[6]) * id;
  hi[7] = (h[1] * h[6] - h[0] * h[7]) * id;
  hi[8] = (h[0] * h[4] - h[1] * h[3]) * id;
  return det;
}

// |G|^2 for integer (a,b,c): G = 2 pi (a hi[0..2], ...) as in set_cp_ewald.C:563-566
double gsq(const double* hi, int a, int b, int c) {
  const double t = 2.0 * M_PI;
  const double gx = (a * hi[0] + b * hi[1] + c * hi[2]) * t;
  const double gy = (a * hi[3] + b * hi[4] + c * hi[5]) * t;
  const double gz = (a * hi[6] + b * hi[7] + c * hi[8]) * t;
  return gx * gx + gy * gy + gz * gz;
}

// Half-sphere enumeration of setkvec3d_sm (set_cp_ewald.C:523-680) with the
// kmax of calc_cutoff (set_cp_ewald.C:176-197), then the kx=0 mirror of
// flip_data_set (para_grp_parse.C:178-285).  The half-sphere test
// |G|^2/2 <= ecut_Ha is |G|^2 <= ecut_ry.
void enumerate_kvecs(double ecut_ry, const double hmat[9], std::vector<int>& kx,
                     std::vector<int>& ky, std::vector<int>& kz) {
  double hi[9];
  inv3(hmat, hi);
  int kmax[3];
  for (int d = 0; d < 3; d++) {
    const double len = std::sqrt(hi[d] * hi[d] + hi[d + 3] * hi[d + 3] + hi[d + 6] * hi[d + 6]);
    const double t = std::sqrt(0.25 * ecut_ry) / (M_PI * len);
    int it = (int)t;
    if (t - it > 0.99) it++;
    kmax[d] = it;
  }
  auto inside = [&](int a, int b, int c) { return gsq(hi, a, b, c) <= ecut_ry; };
  std::vector<int> ha, hb, hc;
  int amax = 0;
  while (amax + 1 <= kmax[0] && inside(amax + 1, 0, 0)) amax++;
  for (int a = 0; a <= amax; a++) {
    int bmin = (a == 0) ? 0 : -kmax[1];
    while (bmin < 0 && !inside(a, bmin, 0)) bmin++;
    int bmax = 0;
    while (bmax + 1 <= kmax[1] && inside(a, bmax + 1, 0)) bmax++;
    for (int b = bmin; b <= bmax; b++) {
      int cmin = (a == 0 && b == 0) ? 1 : -kmax[2];
      while (cmin < 0 && !inside(a, b, cmin)) cmin++;
      int cmax = 0;
      while (cmax + 1 <= kmax[2] && inside(a, b, cmax + 1)) cmax++;
      for (int c = cmin; c <= cmax; c++) {
        ha.push_back(a);
        hb.push_back(b);
        hc.push_back(c);
      }
    }
  }
  // the G=0 term is appended after the half sphere; the mirror of the kx=0
  // half plane is prepended in reverse order, then G=0, then the rest
  int n0 = 0;
  while (n0 < (int)ha.size() && ha[n0] == 0) n0++;
  kx.clear();
  ky.clear();
  kz.clear();
  for (int i = n0 - 1; i >= 0; i--) {
    kx.push_back(0);
    ky.push_back(-hb[i]);
    kz.push_back(-hc[i]);
  }
  kx.push_back(0);
  ky.push_back(0);
  kz.push_back(0);
  for (size_t i = 0; i < ha.size(); i++) {
    kx.push_back(ha[i]);
    ky.push_back(hb[i]);
    kz.push_back(hc[i]);
  }
}

bool make_plan(int n, FftPlan& p) {
  p.n = n;
  p.npass = 0;
  int m = n;
  const int order[4] = {4, 5, 3, 2};
  for (int r : order)
    while (m % r == 0) {
      if (p.npass >= OA_MAXPASS) return false;
      p.radix[p.npass++] = r;
      m /= r;
    }
  return m == 1 && n > 1;
}

int imod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

}  // namespace

struct oa_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  oa_options opt{};
  bool fp32 = false;
  int nx = 0, ny = 0, nz = 0, nst = 0, npts = 0, ld = 0, nlines = 0, nplane_x = 0, nkx0 = 0,
      nred = 0;
  double hmat[9]{}, hmati[9]{}, vol = 0.0, ecut_ry = 0.0;
  std::vector<int> kx, ky, kz;
  // device: the allocations of oa_create come from one block (arena), so a
  // context is released with one hipFree (each hipFree waits for the whole
  // device, stalling every other stream); later ones are separate blocks
  std::vector<void*> allocs;
  char* arena = nullptr;
  size_t arena_size = 0, arena_used = 0;
  char* stage = nullptr;  // staging block of set-up uploads, grown on demand
  size_t stage_cap = 0;
  double2 *psi = nullptr, *psi2 = nullptr, *force = nullptr;
  int *blk_pt0 = nullptr, *pt_code = nullptr, *line_dst = nullptr, *kylo = nullptr,
      *kyhi = nullptr, *red = nullptr, *partner = nullptr,
      *sd_twin = nullptr;  // per point: -1, -2 (redundant kx=0 point) or its redundant twin
  double2 *twx = nullptr, *twy = nullptr, *twz = nullptr;
  double *g2 = nullptr, *occ = nullptr, *dtm = nullptr, *vks = nullptr, *rho = nullptr;
  void *zbuf = nullptr, *xbuf = nullptr, *psir = nullptr, *zout = nullptr;
  void* psk = nullptr;  // plane engine, opt.psir_keep: psi(r) between the plane passes
  double* eke_part = nullptr;
  int nblk = 0;
  double *S = nullptr, *T = nullptr, *lam = nullptr, *gpart = nullptr, *S2 = nullptr;
  int nslice_sym = 0, nslice_full = 0;  // gram K slices (symmetric S, full Lambda)
  double *nsX0 = nullptr, *nsX1 = nullptr, *nsT0 = nullptr, *nsT1 = nullptr, *nsC = nullptr,
         *errpart = nullptr;
  oa::Ns32Bufs ns32{};  // FP32 Newton-Schulz work space (mixed contexts with ns_fp32)
  double *fpart = nullptr, *dpart = nullptr;
  int nfpart = 0, ndpart = 0;
  StepDev* res = nullptr;
  StepDev* res_host = nullptr;
  FftCtx fc{};
  bool timing = false;
  int timing_level = 1;  // 1: every stage, 2: the step and the R->G plane pass
  // ring of per-step event sets so a run of steps can be timed without host
  // synchronisation inside it
  static constexpr int kRing = 64;
  std::vector<hipEvent_t> ev;  // [kRing][OA_NTIMINGS] once timing is enabled
  int ev_cur = 0, ev_count = 0;
  bool have_S = false, have_T = false, have_v = false;
  // Newton-Schulz: 0 two launches per iteration, 1 one cooperative launch
  // over 32 x 32 tiles, 2 one cooperative launch over 16-row blocks
  int ns_kind = 0;
  std::vector<int> ns_kinds;  // the kinds this context can run
  bool sd_fused = false;  // SD update inside the Lambda rotate (nst <= 128)
  int engine = 0;     // kEngFft2, kEngSep, kEngFused or kEngPlane
  // plane engine with V fixed during the step: rho partials come from the
  // R->G pass (psi(r) is recomputed there anyway), no G->R plane pass
  bool fuse_rho = false;
  // graph mode: the two coefficient buffers, and per parity (which buffer
  // holds psi) the chain before Newton-Schulz and the one after it
  double2* psi_buf[2] = {nullptr, nullptr};
  // sizing: device bytes allocated (256-byte granules); dry = count, allocate nothing
  size_t bytes = 0;
  bool dry = false;
  // conjugate gradient: direction [nst][ld] and device state
  double2* cgdir = nullptr;
  oa::CgDev* cgs = nullptr;
  hipGraphExec_t gpre[2] = {nullptr, nullptr}, gpost[2] = {nullptr, nullptr};
  int steps_run = 0;
  int nzblk = 0;      // EKE partials per state of the active z kernel
  int4* lines = nullptr;
  // V_KS from rho (opt.vks_from_rho): the density-sphere box and its buffers
  oa::VksArgs va{};
  double2* vloc = nullptr;
  // autotuner (opt.autotune): candidate launch shapes (buffers are sized for
  // the largest) and what was measured
  std::vector<int> tune_nchunk, tune_nchunk_r, tune_nslice_sym, tune_nslice_full;
  std::vector<int> tune_vks;  // V_KS rows / columns / lines per workgroup
  std::string tune_log;
  // non-local KB pseudopotential (oa_nonlocal_set): projector table, the
  // projections Z, their scaled transpose and the gram partials
  int nl_natm = 0, nl_cap = 0, nl_nslice = 0;  // atoms set, atoms allocated, K slices
  double* nl_sf = nullptr;    // [4 nl_cap][2 ld]: rows of oa::nl_sf_build
  double* nl_z = nullptr;     // [nst][4 nl_cap]
  double* nl_mt = nullptr;    // [nl_cap][nst]
  double* nl_vn = nullptr;    // [nl_cap] vnorm / Omega per atom
  double* nl_part = nullptr;  // [nl_nslice][nst][4 nl_cap]
  double* nl_fatm = nullptr;  // [nl_cap][3]
};

namespace {
constexpr int kEngFft2 = 0;   // two-factor register engine, one kernel per pass
constexpr int kEngSep = 1;    // Stockham LDS kernels, one kernel per pass
constexpr int kEngFused = 2;  // Stockham LDS, y and x passes fused per plane
constexpr int kEngPlane = 3;  // two-factor engine, y and x fused per plane, psi(r) recomputed
constexpr int kMaxRhoChunks = 16;  // rho partials (N doubles each) of one plane pass
}  // namespace

namespace {

oa_status fail(oa_ctx* c, oa_status s, const std::string& m) {
  if (c) c->err = m;
  return s;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(c, OA_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

template <typename P>
oa_status dalloc(oa_ctx* c, P** p, size_t bytes) {
  void* q = nullptr;
  if (bytes == 0) bytes = 16;
  const size_t a = (bytes + 255) & ~(size_t)255;
  c->bytes += a;
  if (c->dry) {  // sizing run: count only
    *p = nullptr;
    return OA_OK;
  }
  if (c->arena && c->arena_used + a <= c->arena_size) {
    *p = reinterpret_cast<P*>(c->arena + c->arena_used);
    c->arena_used += a;
    return OA_OK;
  }
  if (hipMalloc(&q, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, OA_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  c->allocs.push_back(q);
  *p = reinterpret_cast<P*>(q);
  return OA_OK;
}

#define TRY(expr)                     \
  do {                                \
    oa_status s_ = (expr);            \
    if (s_ != OA_OK) return s_;       \
  } while (0)

// the context's staging block, at least `bytes` long (kept until destroy:
// a per-call hipMalloc/hipFree pair would wait for the whole device)
oa_status stage_buf(oa_ctx* c, size_t bytes, char** out) {
  if (bytes > c->stage_cap) {
    if (c->stage) {
      HIPCHK(hipStreamSynchronize(c->stream));
      (void)hipFree(c->stage);
      c->stage = nullptr;
      c->stage_cap = 0;
    }
    if (hipMalloc((void**)&c->stage, bytes) != hipSuccess) {
      (void)hipGetLastError();
      c->stage = nullptr;
      return fail(c, OA_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    c->stage_cap = bytes;
  }
  *out = c->stage;
  return OA_OK;
}

oa_status upload(oa_ctx* c, void* dst, const void* src, size_t bytes) {
  if (c->dry) return OA_OK;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return OA_OK;
}

std::vector<double2> twiddles(int n) {
  std::vector<double2> t(n);
  for (int j = 0; j < n; j++) {
    const double a = 2.0 * M_PI * (double)j / (double)n;
    t[j] = make_double2(std::cos(a), -std::sin(a));
  }
  return t;
}

// States per chunk for a grid of units x nchunk workgroups of which `slots`
// are resident at once: minimise rounds * (states per workgroup + 1), the
// +1 standing for a workgroup's fixed start-up cost; ties go to fewer
// chunks.  At most max_chunks chunks.
void pick_chunks(int units, int slots, int nst, int max_chunks, int& cs, int& nchunk) {
  double best = 1e300;
  cs = nst;
  nchunk = 1;
  for (int nc = 1; nc <= std::min(nst, std::max(1, max_chunks)); nc++) {
    const int s = (nst + nc - 1) / nc, ncr = (nst + s - 1) / s;
    if (ncr != nc) continue;
    const long wg = (long)units * ncr;
    const long rounds = (wg + slots - 1) / slots;
    const double cost = (double)rounds * (s + 1);
    if (cost < best) {
      best = cost;
      cs = s;
      nchunk = ncr;
    }
  }
}

// Autotuner candidates around a heuristic chunk count h: every count in
// [h/2, 2h] (at most max_chunks) that splits nst into equal chunks.
std::vector<int> chunk_candidates(int nst, int h, int max_chunks) {
  std::vector<int> out;
  for (int nc = std::max(1, h / 2); nc <= std::min({nst, max_chunks, 2 * h}); nc++) {
    const int s = (nst + nc - 1) / nc;
    if ((nst + s - 1) / s == nc) out.push_back(nc);
  }
  if (out.empty()) out.push_back(h);
  return out;
}

// Validate the stored layout and build the device tables.
oa_status build_layout(oa_ctx* c) {
  const int n = c->npts;
  const std::vector<int>&kx = c->kx, &ky = c->ky, &kz = c->kz;
  // kx = 0 plane first (stateSlab.C:325-340), kx >= 0 (Gamma half space)
  int nkx0 = 0;
  while (nkx0 < n && kx[nkx0] == 0) nkx0++;
  for (int i = nkx0; i < n; i++)
    if (kx[i] <= 0) return fail(c, OA_EINVAL, "kx=0 points must come first and kx >= 0");
  c->nkx0 = nkx0;
  // lines: runs of constant (kx,ky), kz consecutive (util.C:1364-1397)
  std::vector<int> loff, llen, lkx, lky, lkz0;
  std::map<std::pair<int, int>, int> seen;
  int kxmax = 0;
  for (int i = 0; i < n; i++) {
    if (i == 0 || kx[i] != kx[i - 1] || ky[i] != ky[i - 1]) {
      if (seen.count({kx[i], ky[i]})) return fail(c, OA_EINVAL, "line (kx,ky) split in two runs");
      seen[{kx[i], ky[i]}] = (int)loff.size();
      loff.push_back(i);
      llen.push_back(0);
      lkx.push_back(kx[i]);
      lky.push_back(ky[i]);
      lkz0.push_back(kz[i]);
    } else if (kz[i] != kz[i - 1] + 1) {
      return fail(c, OA_EINVAL, "kz must ascend by one along a line");
    }
    llen.back()++;
    kxmax = std::max(kxmax, kx[i]);
    if (2 * std::abs(ky[i]) >= c->ny || 2 * std::abs(kz[i]) >= c->nz)
      return fail(c, OA_EINVAL, "k-vector outside the FFT grid");
  }
  c->nlines = (int)loff.size();
  c->nplane_x = kxmax + 1;
  if (2 * c->nplane_x > c->nx) return fail(c, OA_EINVAL, "kx range too large for the x grid");
  if (c->nz >= 4096) return fail(c, OA_EINVAL, "nz >= 4096 unsupported");
  // populated ky range per kx column must be contiguous
  std::vector<int> lo(c->nplane_x, 1 << 30), hi(c->nplane_x, -(1 << 30)), cnt(c->nplane_x, 0);
  for (int l = 0; l < c->nlines; l++) {
    lo[lkx[l]] = std::min(lo[lkx[l]], lky[l]);
    hi[lkx[l]] = std::max(hi[lkx[l]], lky[l]);
    cnt[lkx[l]]++;
  }
  int kabs = c->nplane_x - 1;
  for (int x = 0; x < c->nplane_x; x++) {
    if (cnt[x] == 0) {
      lo[x] = 1;
      hi[x] = 0;
    } else if (hi[x] - lo[x] + 1 != cnt[x]) {
      return fail(c, OA_EINVAL, "populated ky range of a kx column is not contiguous");
    } else {
      kabs = std::max(kabs, std::max(std::abs(lo[x]), std::abs(hi[x])));
    }
  }
  // the plane (z) passes skip the transforms' zero inputs and unkept
  // outputs when every populated |kx|, |ky| (|kz|) is <= (N - 1) / 4
  int kzabs = 0;
  for (int i = 0; i < n; i++) kzabs = std::max(kzabs, std::abs(kz[i]));
  c->fc.prune = !c->opt.fft_full && 4 * kabs + 1 <= std::min(c->nx, c->ny);
  c->fc.prune_z = !c->opt.fft_full && 4 * kzabs + 1 <= c->nz;
  // redundant kx=0 points and their partners (CP_State_GSpacePlane.C:2862-3061)
  std::map<std::pair<int, int>, int> kx0idx;
  for (int i = 0; i < nkx0; i++) kx0idx[{ky[i], kz[i]}] = i;
  std::vector<int> red, partner;
  for (int i = 0; i < nkx0; i++) {
    if (ky[i] < 0 || (ky[i] == 0 && kz[i] < 0)) {
      auto it = kx0idx.find({-ky[i], -kz[i]});
      if (it == kx0idx.end()) return fail(c, OA_EINVAL, "redundant kx=0 point without partner");
      red.push_back(i);
      partner.push_back(it->second);
    }
  }
  c->nred = (int)red.size();
  // launch shapes
  FftCtx& f = c->fc;
  if (!make_plan(c->nx, f.px) || !make_plan(c->ny, f.py) || !make_plan(c->nz, f.pz))
    return fail(c, OA_EINVAL, "FFT sizes must be products of 2, 3, 5 (<= 8 passes)");
  f.lpb = std::max(1, 1024 / c->nz);
  {
    const int nch = std::max(1, (c->nplane_x * c->ny + 1023) / 1024);
    f.kc = (c->nplane_x + nch - 1) / nch;
  }
  f.npair = std::max(1, std::min(OA_XSLOTS * OA_NTHR, 512) / c->nx);
  if (f.npair * c->nx > OA_XSLOTS * OA_NTHR) return fail(c, OA_EINVAL, "nx too large");
  // FFT engine.  Mode 0 (auto): the fused two-factor plane passes when the
  // plane is square and its size compiled, else the two-factor engine with
  // one kernel per pass when all three lengths are compiled, else the
  // Stockham kernels.  Modes 1-4 force: 1 Stockham per pass, 2 Stockham fused
  // planes, 3 two-factor per pass, 4 two-factor fused planes.
  const int npairs_all = (c->ny + 1) / 2;
  int lds_max = 0;
  if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) !=
          hipSuccess || lds_max <= 0)
    lds_max = 65536;
  lds_max = std::min(lds_max, 160 * 1024);
  const bool fft2_ok =
      oa::fft2_ntr(c->nx) > 0 && oa::fft2_ntr(c->ny) > 0 && oa::fft2_ntr(c->nz) > 0;
  const size_t p3 = (c->nx == c->ny) ? oa::plane3_smem(c->nx, c->nplane_x, c->fp32) : 0;
  const bool plane_ok = fft2_ok && p3 > 0 && p3 <= (size_t)lds_max;
  const int mode = c->opt.fft_mode;
  if (mode < 0 || mode > 4) return fail(c, OA_EINVAL, "fft_mode must be 0..4");
  if ((mode == 3 && !fft2_ok) || (mode == 4 && !plane_ok))
    return fail(c, OA_EINVAL, "the requested FFT engine is not compiled for this grid");
  if (mode == 4 || (mode == 0 && plane_ok))
    c->engine = kEngPlane;
  else if (mode == 3 || (mode == 0 && fft2_ok))
    c->engine = kEngFft2;
  else if (mode == 2)
    c->engine = kEngFused;
  else
    c->engine = kEngSep;
  f.plane_cfg = oa::plane_config_for(c->nx * c->ny);
  if (c->engine == kEngFused && f.plane_cfg < 0) c->engine = kEngSep;
  if (c->engine == kEngFft2) {
    const int qblk = (npairs_all + oa::fft2_ntr(c->nx) - 1) / oa::fft2_ntr(c->nx);
    const int nchunk = std::max(1, std::min(c->nst, (1024 + qblk * c->nz - 1) / (qblk * c->nz)));
    f.cs = (c->nst + nchunk - 1) / nchunk;
    f.nchunk = (c->nst + f.cs - 1) / f.cs;
    c->nzblk = (c->nlines + oa::fft2_ntr(c->nz) - 1) / oa::fft2_ntr(c->nz);
  }
  if (c->engine == kEngPlane) {
    // State chunks of both plane passes from the resident-workgroup count
    // (every G->R chunk also writes one rho partial of N doubles)
    int rowparts = 1, occ_g = 1, occ_r = 1, rho_ok = 0, ncu = 0;
    if (oa::plane3_info(c->nx, c->nplane_x, c->fp32, f.prune != 0, &rowparts, &occ_g, &occ_r,
                        &rho_ok) !=
        hipSuccess)
      return fail(c, OA_EHIP, "plane pass occupancy query failed");
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) !=
            hipSuccess || ncu <= 0)
      ncu = 256;
    // fused rho: every R->G chunk then writes one rho partial too, so its
    // chunk count is bounded like the G->R one's
    c->fuse_rho = rho_ok && !c->opt.vks_from_rho;
    const int max_r = c->fuse_rho ? std::min(c->nst, kMaxRhoChunks) : c->nst;
    pick_chunks(c->nz * rowparts, std::max(1, occ_g) * ncu, c->nst, kMaxRhoChunks, f.cs,
                f.nchunk);
    pick_chunks(c->nz, std::max(1, occ_r) * ncu, c->nst, max_r, f.cs_r, f.nchunk_r);
    c->nzblk = (c->nlines + oa::fft2_ntr(c->nz) - 1) / oa::fft2_ntr(c->nz);
    if (c->opt.autotune) {
      c->tune_nchunk = chunk_candidates(c->nst, f.nchunk, kMaxRhoChunks);
      c->tune_nchunk_r = chunk_candidates(c->nst, f.nchunk_r, max_r);
    }
  }
  if (c->engine == kEngFused) {
    const int npairs = (c->ny + 1) / 2;
    const int pc0 = std::min(npairs, std::max(1, 2048 / c->nx));
    bool ok = false;
    for (int pc = pc0; pc >= 1 && !ok; pc = (pc > 1) ? pc / 2 : 0) {
      f.pcp = pc;
      f.kcp = std::min(c->nplane_x, std::max(1, (pc * c->nx) / c->ny));
      if ((int)oa::plane_smem(f, c->fp32) <= lds_max) ok = true;
    }
    if (!ok) c->engine = kEngSep;
    const int nchunk = std::max(1, std::min(c->nst, (1024 + c->nz - 1) / c->nz));
    f.cs = (c->nst + nchunk - 1) / nchunk;
    f.nchunk = (c->nst + f.cs - 1) / f.cs;
  }
  c->nblk = (c->nlines + f.lpb - 1) / f.lpb;
  if (c->engine != kEngFft2 && c->engine != kEngPlane) c->nzblk = c->nblk;
  std::vector<int> blk_pt0(c->nblk + 1), pt_code(n), line_dst(c->nlines);
  std::vector<int4> lines(c->nlines);
  for (int b = 0; b < c->nblk; b++) blk_pt0[b] = loff[b * f.lpb];
  blk_pt0[c->nblk] = n;
  for (int l = 0; l < c->nlines; l++) {
    line_dst[l] = lkx[l] * c->ny + imod(lky[l], c->ny);
    lines[l] = make_int4(loff[l], llen[l], lkz0[l], line_dst[l]);
    const int li = l % f.lpb;
    for (int i = 0; i < llen[l]; i++) {
      const int p = loff[l] + i;
      const unsigned code = ((lkx[l] == 0) ? 0x80000000u : 0u) | ((unsigned)li << 12) |
                            (unsigned)imod(lkz0[l] + i, c->nz);
      pt_code[p] = (int)code;
    }
  }
  std::vector<double> g2(n), dtm(n);
  for (int i = 0; i < n; i++) g2[i] = gsq(c->hmati, kx[i], ky[i], kz[i]);
  // dt / cmass with CP_create_mass (cp_control_integrate.C:136-203)
  {
    const double tau = c->opt.tau_mass / kTimeConv;
    const double cmass0 = 2.0 * tau * tau * kCpEmagic;
    const double ecm = c->opt.ecut_mass;
    for (int i = 0; i < n; i++) {
      const double w = (kx[i] == 0) ? 1.0 : 2.0;
      const double m = (g2[i] > ecm) ? (w * g2[i]) * (cmass0 / ecm) : w * cmass0;
      dtm[i] = c->opt.dt / m;
    }
  }
  TRY(dalloc(c, &c->blk_pt0, sizeof(int) * blk_pt0.size()));
  TRY(dalloc(c, &c->pt_code, sizeof(int) * n));
  TRY(dalloc(c, &c->line_dst, sizeof(int) * c->nlines));
  TRY(dalloc(c, &c->kylo, sizeof(int) * c->nplane_x));
  TRY(dalloc(c, &c->kyhi, sizeof(int) * c->nplane_x));
  TRY(dalloc(c, &c->red, sizeof(int) * std::max(1, c->nred)));
  TRY(dalloc(c, &c->partner, sizeof(int) * std::max(1, c->nred)));
  TRY(dalloc(c, &c->sd_twin, sizeof(int) * n));
  TRY(dalloc(c, &c->g2, sizeof(double) * n));
  TRY(dalloc(c, &c->dtm, sizeof(double) * n));
  TRY(upload(c, c->blk_pt0, blk_pt0.data(), sizeof(int) * blk_pt0.size()));
  TRY(upload(c, c->pt_code, pt_code.data(), sizeof(int) * n));
  TRY(upload(c, c->line_dst, line_dst.data(), sizeof(int) * c->nlines));
  TRY(dalloc(c, &c->lines, sizeof(int4) * c->nlines));
  TRY(upload(c, c->lines, lines.data(), sizeof(int4) * c->nlines));
  TRY(upload(c, c->kylo, lo.data(), sizeof(int) * c->nplane_x));
  TRY(upload(c, c->kyhi, hi.data(), sizeof(int) * c->nplane_x));
  if (c->nred) {
    TRY(upload(c, c->red, red.data(), sizeof(int) * c->nred));
    TRY(upload(c, c->partner, partner.data(), sizeof(int) * c->nred));
  }
  {  // the SD update writes each redundant point with its partner
    std::vector<int> twin(n, -1);
    for (int r = 0; r < c->nred; r++) {
      twin[partner[r]] = red[r];
      twin[red[r]] = -2;
    }
    TRY(upload(c, c->sd_twin, twin.data(), sizeof(int) * n));
  }
  TRY(upload(c, c->g2, g2.data(), sizeof(double) * n));
  TRY(upload(c, c->dtm, dtm.data(), sizeof(double) * n));
  auto tx = twiddles(c->nx), ty = twiddles(c->ny), tz = twiddles(c->nz);
  TRY(dalloc(c, &c->twx, sizeof(double2) * c->nx));
  TRY(dalloc(c, &c->twy, sizeof(double2) * c->ny));
  TRY(dalloc(c, &c->twz, sizeof(double2) * c->nz));
  TRY(upload(c, c->twx, tx.data(), sizeof(double2) * c->nx));
  TRY(upload(c, c->twy, ty.data(), sizeof(double2) * c->ny));
  TRY(upload(c, c->twz, tz.data(), sizeof(double2) * c->nz));
  return OA_OK;
}

// The density sphere |G|^2 <= 4 ecut_ry (cp_hart_ext.C:111, 175) and its box
// kx in [0, kxr], |ky| <= kyr, |kz| <= kzr; the grid must hold it without
// aliasing (2 k + 1 <= N on every axis).
oa_status alloc_vks(oa_ctx* c) {
  oa::VksArgs& a = c->va;
  const double ecut4 = 4.0 * c->ecut_ry;
  int kxr = 0, kyr = 0, kzr = 0;
  for (int i = 0; i <= c->nx / 2; i++)
    for (int j = -c->ny / 2; j <= c->ny / 2; j++)
      for (int k = -c->nz / 2; k <= c->nz / 2; k++)
        if (gsq(c->hmati, i, j, k) <= ecut4) {
          kxr = std::max(kxr, i);
          kyr = std::max(kyr, std::abs(j));
          kzr = std::max(kzr, std::abs(k));
        }
  if (2 * kxr + 1 > c->nx || 2 * kyr + 1 > c->ny || 2 * kzr + 1 > c->nz)
    return fail(c, OA_EINVAL,
                "grid too small for the density sphere |G|^2 <= 4 ecut (needs N >= " +
                    std::to_string(2 * kxr + 1) + ", " + std::to_string(2 * kyr + 1) + ", " +
                    std::to_string(2 * kzr + 1) + ")");
  a.nx = c->nx;
  a.ny = c->ny;
  a.nz = c->nz;
  a.nkx = kxr + 1;
  a.kyr = kyr;
  a.nky = 2 * kyr + 1;
  a.pr = 8;
  a.cb = 8;
  a.lb = 8;
  a.px = c->fc.px;
  a.py = c->fc.py;
  a.pz = c->fc.pz;
  std::memcpy(a.hmati, c->hmati, sizeof(a.hmati));
  a.vol = c->vol;
  a.ecut4 = ecut4;
  if (oa::vks_smem(a) > 64 * 1024)
    return fail(c, OA_EINVAL, "V_KS pass needs more than 64 KiB of LDS");
  // autotune: row pairs (x passes), columns (y passes) and z lines per
  // workgroup among 2, 4, 8, 16 where the LDS fits; the partial-sum arrays
  // are sized for the most workgroups
  int vmin = a.pr;
  if (c->opt.autotune) {
    for (int v : {2, 4, 8, 16}) {
      oa::VksArgs t = a;
      t.pr = t.cb = t.lb = v;
      if (oa::vks_smem(t) <= 64 * 1024) c->tune_vks.push_back(v);
    }
    vmin = c->tune_vks.front();
  }
  oa::VksArgs amin = a;
  amin.pr = amin.cb = amin.lb = vmin;
  const size_t nbox = (size_t)a.nkx * a.nky * c->nz;
  TRY(dalloc(c, &a.hx, (size_t)c->nz * a.nkx * c->ny * sizeof(double2)));
  TRY(dalloc(c, &a.hl, nbox * sizeof(double2)));
  TRY(dalloc(c, &c->vloc, nbox * sizeof(double2)));
  TRY(dalloc(c, &a.part_x, (size_t)oa::vks_nbx(amin) * 3 * sizeof(double)));
  TRY(dalloc(c, &a.part_l, (size_t)oa::vks_nbl(amin) * 2 * sizeof(double)));
  a.vloc = c->vloc;
  return OA_OK;
}

oa_status alloc_state(oa_ctx* c, const double* occ) {
  const size_t N = (size_t)c->nx * c->ny * c->nz;
  const size_t gsz = (size_t)c->nst * c->ld * sizeof(double2);
  const size_t rsz = c->fp32 ? 4 : 8;
  const size_t csz = 2 * rsz;
  TRY(dalloc(c, &c->psi, gsz));
  TRY(dalloc(c, &c->psi2, gsz));
  c->psi_buf[0] = c->psi;
  c->psi_buf[1] = c->psi2;
  if (c->opt.use_cg) {
    TRY(dalloc(c, &c->cgdir, gsz));
    TRY(dalloc(c, &c->cgs, sizeof(oa::CgDev)));
    if (!c->dry) {
      HIPCHK(oa::fill_zero(c->cgdir, gsz, c->stream));
      HIPCHK(oa::fill_zero(c->cgs, sizeof(oa::CgDev), c->stream));
    }
  }
  TRY(dalloc(c, &c->force, gsz));
  const size_t zsz = (size_t)c->nst * c->nz * c->nplane_x * c->ny * csz;
  TRY(dalloc(c, &c->zbuf, zsz));
  if (c->engine != kEngSep) {
    int np = std::max(c->fc.nchunk, c->tune_nchunk.empty() ? 0 : c->tune_nchunk.back());
    if (c->fuse_rho)
      np = std::max({np, c->fc.nchunk_r, c->tune_nchunk_r.empty() ? 0 : c->tune_nchunk_r.back()});
    TRY(dalloc(c, &c->fc.rho_part, (size_t)np * N * sizeof(double)));
  }
  if (c->engine == kEngPlane) {
    // one state's psi(r) for oa_psir_get; every state's when the passes
    // run separately and opt.psir_keep asks to keep it between them
    TRY(dalloc(c, &c->zout, zsz));
    TRY(dalloc(c, &c->psir, N * rsz));
    if (c->opt.psir_keep && !c->fuse_rho) {
      const size_t kper = oa::plane3_kper(c->nx, c->nplane_x, c->fp32);
      if (kper == 0) return fail(c, OA_EINVAL, "plane size not compiled");
      TRY(dalloc(c, &c->psk, (size_t)c->nst * c->nz * kper * csz));
    }
  } else {
    if (c->engine != kEngFused)
      TRY(dalloc(c, &c->xbuf, (size_t)c->nst * c->nz * c->ny * c->nplane_x * csz));
    TRY(dalloc(c, &c->psir, (size_t)c->nst * N * rsz));
  }
  TRY(dalloc(c, &c->rho, N * sizeof(double)));
  TRY(dalloc(c, &c->vks, N * sizeof(double)));
  TRY(dalloc(c, &c->occ, c->nst * sizeof(double)));
  TRY(dalloc(c, &c->eke_part, (size_t)c->nst * c->nzblk * sizeof(double)));
  const int n = c->nst;
  const size_t nn = (size_t)n * n;
  TRY(dalloc(c, &c->S, nn * sizeof(double)));
  TRY(dalloc(c, &c->S2, nn * sizeof(double)));
  TRY(dalloc(c, &c->T, nn * sizeof(double)));
  TRY(dalloc(c, &c->lam, nn * sizeof(double)));
  // gram: one workgroup per (128 x 128 output block, K slice), about one
  // workgroup per CU in all; the same slice count serves S and Lambda
  const int K = 2 * c->npts;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) !=
          hipSuccess || ncu <= 0)
    ncu = 256;
  c->nslice_sym = oa::gram_slices(n, K, 2 * c->nkx0, 1, ncu);
  c->nslice_full = oa::gram_slices(n, K, 2 * c->nkx0, 0, ncu);
  int max_slice = std::max(c->nslice_sym, c->nslice_full);
  if (c->opt.autotune) {
    // K slices: half, the heuristic and twice it (at least 2, at most K/64)
    auto cands = [&](int h) {
      std::vector<int> v;
      for (int x : {h / 2, h, 2 * h})
        if (x >= 2 && x <= std::max(2, K / 64) && (v.empty() || v.back() != x)) v.push_back(x);
      if (v.empty()) v.push_back(h);
      return v;
    };
    c->tune_nslice_sym = cands(c->nslice_sym);
    c->tune_nslice_full = cands(c->nslice_full);
    max_slice = std::max({max_slice, c->tune_nslice_sym.back(), c->tune_nslice_full.back()});
  }
  TRY(dalloc(c, &c->gpart, (size_t)max_slice * nn * sizeof(double)));
  TRY(dalloc(c, &c->nsX0, nn * sizeof(double)));
  TRY(dalloc(c, &c->nsX1, nn * sizeof(double)));
  TRY(dalloc(c, &c->nsT0, nn * sizeof(double)));
  TRY(dalloc(c, &c->nsT1, nn * sizeof(double)));
  TRY(dalloc(c, &c->nsC, nn * sizeof(double)));
  if (c->fp32 && c->opt.ns_fp32) {
    TRY(dalloc(c, &c->ns32.X0, nn * sizeof(float)));
    TRY(dalloc(c, &c->ns32.X1, nn * sizeof(float)));
    TRY(dalloc(c, &c->ns32.T0, nn * sizeof(float)));
    TRY(dalloc(c, &c->ns32.T1, nn * sizeof(float)));
    TRY(dalloc(c, &c->ns32.C, nn * sizeof(float)));
  }
  const int nt = (n + 31) / 32;
  // per-workgroup error partials of either cooperative kind
  TRY(dalloc(c, &c->errpart,
             (size_t)(c->opt.ns_max_iter + 2) *
                 std::max(nt * nt, oa::kNsRowsMaxSplit * ((n + 15) / 16)) * sizeof(double)));
  c->ns_kinds.clear();
  const int nsm = c->opt.ns_mode;
  if (nsm == 0) {
    // row blocks split over 4, 2 or 1 workgroups (kinds 4, 3, 2)
    for (int kind = 4; kind >= 2; kind--)
      if (!c->dry && oa::ns_rows_possible(n, oa::ns_rows_split(kind), c->device))
        c->ns_kinds.push_back(kind);
    if (!c->dry && oa::ns_coop_possible(n, c->device)) c->ns_kinds.push_back(1);
    if (n <= oa::kNsOneMax) c->ns_kinds.push_back(oa::kNsOneKind);
    c->ns_kinds.push_back(0);
  } else if (nsm == 1) {
    c->ns_kinds.push_back(0);
  } else if (nsm == 6) {  // no cooperative launch
    c->ns_kinds.push_back(n <= oa::kNsOneMax ? oa::kNsOneKind : 0);
  } else if (nsm == 2) {
    if (!c->dry && !oa::ns_coop_possible(n, c->device))
      return fail(c, OA_EINVAL, "ns_mode 2: the 32 x 32 tiles of n do not fit on the device at once");
    c->ns_kinds.push_back(1);
  } else if (nsm >= 3 && nsm <= 5) {
    const int kind = nsm - 1;
    if (!c->dry && !oa::ns_rows_possible(n, oa::ns_rows_split(kind), c->device))
      return fail(c, OA_EINVAL,
                  "ns_mode " + std::to_string(nsm) +
                      ": the row blocks of n (16 rows each, split over that many workgroups "
                      "by 16-column tiles) do not fit on the device at once");
    c->ns_kinds.push_back(kind);
  } else {
    return fail(c, OA_EINVAL, "ns_mode must be 0..6");
  }
  c->ns_kind = c->ns_kinds.front();
  c->sd_fused = false;  // measured slower than a separate SD pass (registers)
  // |F|^2 partials: one per workgroup of the SD kernel or of the rotate that
  // applies Lambda and the SD update together
  c->nfpart = std::max(1024, oa::rotate_nwg(n, K));
  c->ndpart = 1024;
  TRY(dalloc(c, &c->fpart, c->nfpart * sizeof(double)));
  TRY(dalloc(c, &c->dpart, c->ndpart * sizeof(double)));
  TRY(dalloc(c, &c->res, sizeof(StepDev)));
  if (c->opt.vks_from_rho) TRY(alloc_vks(c));
  if (c->dry) return OA_OK;  // sizing only
  HIPCHK(hipHostMalloc((void**)&c->res_host, sizeof(StepDev), hipHostMallocDefault));
  std::memset(c->res_host, 0, sizeof(StepDev));
  HIPCHK(oa::fill_zero(c->psi, gsz, c->stream));
  HIPCHK(oa::fill_zero(c->psi2, gsz, c->stream));
  HIPCHK(oa::fill_zero(c->force, gsz, c->stream));
  HIPCHK(oa::fill_zero(c->vks, N * sizeof(double), c->stream));
  HIPCHK(oa::fill_zero(c->res, sizeof(StepDev), c->stream));
  if (c->vloc)
    HIPCHK(oa::fill_zero(c->vloc, (size_t)c->va.nkx * c->va.nky * c->nz * sizeof(double2),
                         c->stream));
  std::vector<double> o(c->nst, 1.0);
  if (occ)
    for (int i = 0; i < c->nst; i++) o[i] = occ[i];
  TRY(upload(c, c->occ, o.data(), sizeof(double) * c->nst));
  // FFT context
  FftCtx& f = c->fc;
  f.g = {c->nx, c->ny, c->nz, c->nplane_x, c->nlines, c->npts, c->ld, c->nst};
  f.blk_pt0 = c->blk_pt0;
  f.pt_code = c->pt_code;
  f.line_dst = c->line_dst;
  f.kylo = c->kylo;
  f.kyhi = c->kyhi;
  f.twx = c->twx;
  f.twy = c->twy;
  f.twz = c->twz;
  f.g2 = c->g2;
  f.occ = c->occ;
  f.vks = c->vks;
  f.zbuf = c->zbuf;
  f.zout = c->zout;
  f.xbuf = c->xbuf;
  f.psir = c->psir;
  f.psk = c->psk;
  f.rho = c->rho;
  f.eke_part = c->eke_part;
  f.lines = c->lines;
  f.inv_vol = 1.0 / c->vol;
  f.ecut_ry = c->ecut_ry;
  f.scale = -2.0 / ((double)c->nx * c->ny * c->nz);
  HIPCHK(hipStreamSynchronize(c->stream));
  return OA_OK;
}

FftCtx& fctx(oa_ctx* c) {
  c->fc.psi = c->psi;
  c->fc.force = c->force;
  return c->fc;
}

// The plane engine's view of the states [s0, s0 + ns): every per-state array
// offset to state s0, the state chunks of both plane passes recounted for ns
// states (the rho partial buffer holds the full count's chunks).
FftCtx range_ctx(oa_ctx* c, int s0, int ns) {
  FftCtx f = fctx(c);
  const size_t zst = (size_t)c->nz * c->nplane_x * c->ny * (c->fp32 ? 8 : 16);
  f.g.nst = ns;
  f.psi = c->psi + (size_t)s0 * c->ld;
  f.force = c->force + (size_t)s0 * c->ld;
  f.zbuf = static_cast<unsigned char*>(c->zbuf) + (size_t)s0 * zst;
  f.zout = static_cast<unsigned char*>(c->zout) + (size_t)s0 * zst;
  if (c->psk)
    f.psk = static_cast<unsigned char*>(c->psk) +
            (size_t)s0 * c->nz * oa::plane3_kper(c->nx, c->nplane_x, c->fp32) * (c->fp32 ? 8 : 16);
  f.occ = c->occ + s0;
  f.eke_part = c->eke_part + (size_t)s0 * c->nzblk;
  f.nchunk = (ns + f.cs - 1) / f.cs;
  f.nchunk_r = (ns + f.cs_r - 1) / f.cs_r;
  return f;
}

oa_status check_range(oa_ctx* c, int s0, int ns) {
  if (s0 < 0 || ns < 1 || s0 > c->nst - ns)
    return fail(c, OA_EINVAL, "state range [s0, s0 + ns) outside [0, nstates)");
  if (c->engine != kEngPlane)
    return fail(c, OA_EINVAL,
                "state ranges need the fused plane engine (fft_mode 0 or 4, square x-y planes "
                "of a compiled length)");
  return OA_OK;
}

// Timing marks inside oa_step: ev[0] start, ev[1..6] after each FFT stage,
// ev[7] reductions + Lambda, ev[8] SD + sync, ev[9] S, ev[10] NS,
// ev[11] Psi T, ev[12] residual (end).
// Timing level 2 records only the step's ends and the R->G plane pass
// (marks 0, 3, 4, 12): each recorded event costs the stream a few
// microseconds, which level 1 pays at every stage boundary.
bool mark_on(const oa_ctx* c, int i) {
  return c->timing_level == 1 || i == 0 || i == 3 || i == 4 || i == 12;
}

void mark(oa_ctx* c, int i) {
  if (c->timing && i >= 0 && mark_on(c, i))
    (void)hipEventRecord(c->ev[(size_t)c->ev_cur * OA_NTIMINGS + i], c->stream);
}

// ---- stages (enqueue only) ----
// V_KS[rho] into vks, and E_H, E_xc, E_ext into the step results
oa_status st_vks(oa_ctx* c) {
  oa::VksArgs a = c->va;
  a.rho = c->rho;
  a.v = c->vks;
  a.twx = c->twx;
  a.twy = c->twy;
  a.twz = c->twz;
  HIPCHK(oa::vks_build(a, c->res, c->stream));
  return OA_OK;
}

// Two-factor engine: z, y, x c2r (rho partials), then the rho reduction with
// partial sums of V rho.  Fused Stockham: z, plane forward, rho reduction.
// Separate Stockham: z, y, x c2r + rho.  Marks ev0+0..2.
oa_status st_g2r(oa_ctx* c, int ev0 = -1) {
  auto mk = [&](int k) { mark(c, ev0 < 0 ? -1 : ev0 + k); };
  if (c->engine == kEngPlane) {
    HIPCHK(oa::fft2_launch(oa::FFT_ZFWD, c->nz, c->fp32, fctx(c), c->stream));
    mk(0);
    HIPCHK(oa::plane3_launch(0, c->nx, c->fp32, fctx(c), c->stream));
    mk(1);
    HIPCHK(oa::rho_reduce(fctx(c), c->dpart, c->ndpart, c->stream, !c->opt.vks_from_rho));
    if (c->opt.vks_from_rho) TRY(st_vks(c));
    mk(2);
  } else if (c->engine == kEngFft2) {
    HIPCHK(oa::fft2_launch(oa::FFT_ZFWD, c->nz, c->fp32, fctx(c), c->stream));
    mk(0);
    HIPCHK(oa::fft2_launch(oa::FFT_YFWD, c->ny, c->fp32, fctx(c), c->stream));
    mk(1);
    HIPCHK(oa::fft2_launch(oa::FFT_XC2R, c->nx, c->fp32, fctx(c), c->stream));
    HIPCHK(oa::rho_reduce(fctx(c), c->dpart, c->ndpart, c->stream, !c->opt.vks_from_rho));
    if (c->opt.vks_from_rho) TRY(st_vks(c));
    mk(2);
  } else if (c->engine == kEngFused) {
    HIPCHK(oa::fft_stage(fctx(c), c->fp32, oa::FFT_ZFWD, c->stream));
    mk(0);
    HIPCHK(oa::plane_fwd(fctx(c), c->fp32, c->stream));
    mk(1);
    HIPCHK(oa::rho_reduce(fctx(c), c->dpart, c->ndpart, c->stream, !c->opt.vks_from_rho));
    if (c->opt.vks_from_rho) TRY(st_vks(c));
    mk(2);
  } else {
    for (int s = oa::FFT_ZFWD; s <= oa::FFT_XC2R; s++) {
      HIPCHK(oa::fft_stage(fctx(c), c->fp32, s, c->stream));
      if (s < oa::FFT_XC2R) mk(s);
    }
    if (c->opt.vks_from_rho) TRY(st_vks(c));
    mk(2);
  }
  return OA_OK;
}

// The EKE and E_loc reductions at the end of the R->G transform.
oa_status st_energies(oa_ctx* c) {
  const size_t N = (size_t)c->nx * c->ny * c->nz;
  HIPCHK(oa::reduce_to(c->eke_part, c->nst * c->nzblk, 1.0, &c->res->eke, c->stream));
  if (!c->opt.vks_from_rho)  // else E_ext came with V_KS
    HIPCHK(oa::reduce_to(c->dpart, c->ndpart, c->vol / (double)N, &c->res->eloc, c->stream));
  return OA_OK;
}

// Plane engine, fused rho (V fixed during the step): z forward, the R->G
// plane pass with the rho partials of its chunks, the rho reduction, z
// back.  Marks ev0+0..5 (1 and 2 empty), then the energy reductions.
oa_status st_g2r2g_rho(oa_ctx* c, int ev0 = -1) {
  auto mk = [&](int k) { mark(c, ev0 < 0 ? -1 : ev0 + k); };
  HIPCHK(oa::fft2_launch(oa::FFT_ZFWD, c->nz, c->fp32, fctx(c), c->stream));
  mk(0);
  mk(1);
  mk(2);
  FftCtx f = fctx(c);
  HIPCHK(oa::plane3_launch(3, c->nx, c->fp32, f, c->stream));
  mk(3);
  f.nchunk = f.nchunk_r;  // the partials are the R->G chunks'
  HIPCHK(oa::rho_reduce(f, c->dpart, c->ndpart, c->stream));
  mk(4);
  f.zbuf = c->zout;
  HIPCHK(oa::fft2_launch(oa::FFT_ZBWD, c->nz, c->fp32, f, c->stream));
  mk(5);
  return st_energies(c);
}

// Two-factor / separate: x r2c, y, z.  Fused: plane backward, z.
// Marks ev0+3..5, then the EKE and E_loc reductions.
oa_status st_r2g(oa_ctx* c, int ev0 = -1) {
  auto mk = [&](int k) { mark(c, ev0 < 0 ? -1 : ev0 + k); };
  const size_t N = (size_t)c->nx * c->ny * c->nz;
  if (c->engine == kEngPlane) {
    HIPCHK(oa::plane3_launch(1, c->nx, c->fp32, fctx(c), c->stream));
    mk(3);
    mk(4);
    FftCtx f = fctx(c);
    f.zbuf = c->zout;  // the z pass back reads the R->G plane output
    HIPCHK(oa::fft2_launch(oa::FFT_ZBWD, c->nz, c->fp32, f, c->stream));
    mk(5);
  } else if (c->engine == kEngFft2) {
    HIPCHK(oa::fft2_launch(oa::FFT_XR2C, c->nx, c->fp32, fctx(c), c->stream));
    mk(3);
    HIPCHK(oa::fft2_launch(oa::FFT_YBWD, c->ny, c->fp32, fctx(c), c->stream));
    mk(4);
    HIPCHK(oa::fft2_launch(oa::FFT_ZBWD, c->nz, c->fp32, fctx(c), c->stream));
    mk(5);
  } else if (c->engine == kEngFused) {
    HIPCHK(oa::plane_bwd(fctx(c), c->fp32, c->stream));
    mk(3);
    mk(4);
    HIPCHK(oa::fft_stage(fctx(c), c->fp32, oa::FFT_ZBWD, c->stream));
    mk(5);
  } else {
    for (int s = oa::FFT_XR2C; s <= oa::FFT_ZBWD; s++) {
      HIPCHK(oa::fft_stage(fctx(c), c->fp32, s, c->stream));
      mk(s);
    }
    HIPCHK(oa::dot_parts(c->vks, c->rho, N, c->dpart, c->ndpart, c->stream));
  }
  return st_energies(c);
}

// Split-BF16 products of a mixed-precision context (opt.gemm_split bits):
// the Lambda gram, S, Psi T, the Lambda projection F -= w Psi Lambda.  The
// projection and the rotation have a split kernel up to 128 states.
constexpr int kSplitLambda = 1, kSplitS = 2, kSplitRot = 4, kSplitProj = 8;
bool split_on(const oa_ctx* c, int bit) { return c->fp32 && (c->opt.gemm_split & bit) != 0; }
bool split_rotate_ok(const oa_ctx* c) { return c->nst <= 128 && (2 * c->ld) % 64 == 0; }

oa_status st_gram(oa_ctx* c, const double2* A, const double2* B, bool sym, double wlo, double whi,
                  double* C, bool split = false) {
  oa::GramArgs g;
  g.A = reinterpret_cast<const double*>(A);
  g.B = reinterpret_cast<const double*>(B);
  g.n = c->nst;
  g.lda = 2 * c->ld;
  g.K = 2 * c->npts;
  g.k_lo = 2 * c->nkx0;
  g.w_lo = wlo;
  g.w_hi = whi;
  g.sym = sym ? 1 : 0;
  g.nslice = sym ? c->nslice_sym : c->nslice_full;
  g.part = c->gpart;
  g.nb = c->nst;
  if (split) {  // FP32 partials in the same buffer
    float* pf = reinterpret_cast<float*>(c->gpart);
    HIPCHK(oa::gram_bf16(g, pf, c->stream));
    HIPCHK(oa::gram_reduce_f32(pf, g.nslice, c->nst, c->nst, g.sym, 1.0, C, c->stream));
    return OA_OK;
  }
  HIPCHK(oa::gram_f64(g, c->stream));
  HIPCHK(oa::gram_reduce(c->gpart, g.nslice, c->nst, c->nst, g.sym, 1.0, C, c->stream));
  return OA_OK;
}

// sd: also apply psi += dt/m F to X (= psi) in the same pass (n <= 128)
// fnorm: also leave per-workgroup partials of sum Y^2 in fpart (n <= 128)
// split: the split-BF16 kernel (no SD epilogue; split_rotate_ok);
// update: Y = X + X (M - I) (split only)
oa_status st_rotate(oa_ctx* c, const double2* X, const double* M, double2* Y, double beta,
                    double slo, double shi, bool sd = false, bool fnorm = false,
                    bool split = false, bool update = false) {
  oa::RotArgs r{};
  r.X = reinterpret_cast<const double*>(X);
  r.M = M;
  r.Y = reinterpret_cast<double*>(Y);
  r.n = c->nst;
  r.ldx = 2 * c->ld;
  r.K = 2 * c->npts;
  r.k_lo = 2 * c->nkx0;
  r.s_lo = slo;
  r.s_hi = shi;
  r.beta = beta;
  if (sd) {
    r.sd_psi = reinterpret_cast<double*>(const_cast<double2*>(X));
    r.dtm = c->dtm;
  }
  if (sd || fnorm) r.fpart = c->fpart;
  if (split) {
    if (sd) return fail(c, OA_EINVAL, "split rotate without SD epilogue");
    if (update) {
      r.Xadd = reinterpret_cast<const double*>(X);
      r.m_minus_i = 1;
    }
    HIPCHK(oa::rotate_bf16(r, c->stream));
    return OA_OK;
  }
  HIPCHK(oa::rotate_f64(r, c->stream));
  return OA_OK;
}

// workgroups (|F|^2 partials) of the Lambda projection with fnorm
int lambda_nwg(const oa_ctx* c) {
  return split_on(c, kSplitProj) && split_rotate_ok(c) ? oa::rotate_bf16_nwg(2 * c->npts)
                                                           : oa::rotate_nwg(c->nst, 2 * c->npts);
}

// Lambda_ij = sum_stored Re conj(psi_i) F_j (alpha = 1, cpaimd.C:2527) and
// F_j -= w sum_i psi_i Lambda_ij, w = 1/2 on kx=0
// (CP_State_GSpacePlane.C:2076-2086).  with_sd: the same pass also applies
// the SD update and leaves the |F|^2 partials (then st_sd_finish).
oa_status st_lambda(oa_ctx* c, bool with_sd = false) {
  TRY(st_gram(c, c->psi, c->force, false, 1.0, 1.0, c->lam, split_on(c, kSplitLambda)));
  const bool spr = split_on(c, kSplitProj) && split_rotate_ok(c) && !with_sd;
  TRY(st_rotate(c, c->psi, c->lam, c->force, 1.0, -0.5, -1.0, with_sd, false, spr));
  return OA_OK;
}

// |F|^2 from the partials of the fused rotate (FP64 kernel), then the
// redundancy sync
oa_status st_sd_finish(oa_ctx* c) {
  HIPCHK(oa::reduce_to(c->fpart, oa::rotate_nwg(c->nst, 2 * c->npts), 1.0, &c->res->fnorm2,
                       c->stream));
  HIPCHK(oa::red_sync(c->psi, c->red, c->partner, c->nred, c->nst, c->ld, c->stream));
  return OA_OK;
}

// Lambda projection, then the CG direction and update: |F|^2 (fovlap) must
// be complete before gamma, so the update is a pass of its own
oa_status st_lambda_cg(oa_ctx* c) {
  const size_t N2 = (size_t)2 * c->nst * c->ld;  // reals in F, zero padding included
  TRY(st_gram(c, c->psi, c->force, false, 1.0, 1.0, c->lam, split_on(c, kSplitLambda)));
  if (c->nst <= 128) {
    TRY(st_rotate(c, c->psi, c->lam, c->force, 1.0, -0.5, -1.0, false, true,
                  split_on(c, kSplitProj) && split_rotate_ok(c)));
    HIPCHK(oa::reduce_to(c->fpart, lambda_nwg(c), 1.0, &c->res->fnorm2, c->stream));
  } else {
    TRY(st_rotate(c, c->psi, c->lam, c->force, 1.0, -0.5, -1.0));
    const double* fr = reinterpret_cast<const double*>(c->force);
    HIPCHK(oa::dot_parts(fr, fr, N2, c->fpart, c->nfpart, c->stream));
    HIPCHK(oa::reduce_to(c->fpart, c->nfpart, 1.0, &c->res->fnorm2, c->stream));
  }
  HIPCHK(oa::cg_gamma(c->res, c->cgs, c->stream));
  HIPCHK(oa::cg_update(c->psi, c->cgdir, c->force, c->dtm, c->g2, c->ecut_ry, c->nst, c->npts,
                       c->ld, c->cgs, c->stream));
  HIPCHK(oa::red_sync(c->psi, c->red, c->partner, c->nred, c->nst, c->ld, c->stream));
  return OA_OK;
}

// Non-local KB projections of psi: Z = Psi SF^T over the stored points with
// w = 1 on kx = 0 and 2 elsewhere (CP_enl_matrix_calc, cp_nlmat.C:86-118);
// derivs: also the three derivative row groups (nb = 4 natoms).  Leaves
// Z in nl_z ([nst][nb]).
oa_status st_nl_z(oa_ctx* c, bool derivs) {
  oa::GramArgs g;
  g.A = reinterpret_cast<const double*>(c->psi);
  g.B = c->nl_sf;
  g.n = c->nst;
  g.nb = derivs ? 4 * c->nl_natm : c->nl_natm;
  g.lda = 2 * c->ld;
  g.K = 2 * c->npts;
  g.k_lo = 2 * c->nkx0;
  g.w_lo = 1.0;
  g.w_hi = 2.0;
  g.sym = 0;
  g.nslice = c->nl_nslice;
  g.part = c->nl_part;
  HIPCHK(oa::gram_f64(g, c->stream));
  HIPCHK(oa::gram_reduce(c->nl_part, g.nslice, c->nst, g.nb, 0, 1.0, c->nl_z, c->stream));
  return OA_OK;
}

// The step's non-local part: Z, E_NL and Z' = 2 vnorm/Omega Z
// (cp_nl_energy_forc.C:196-216), then F -= w Z' SF
// (CP_enl_force_calc, cp_nl_energy_forc.C:61-85), no occupation weight on
// the force, as in the reference.
oa_status st_nonlocal(oa_ctx* c) {
  TRY(st_nl_z(c, false));
  HIPCHK(oa::nl_energy(c->nl_z, c->nl_natm, c->nst, c->nl_natm, c->nl_vn, c->occ, c->nl_mt,
                       &c->res->enl, c->stream));
  oa::RotArgs r{};
  r.X = c->nl_sf;
  r.M = c->nl_mt;
  r.Y = reinterpret_cast<double*>(c->force);
  r.n = c->nst;
  r.ni = c->nl_natm;
  r.ldx = 2 * c->ld;
  r.K = 2 * c->npts;
  r.k_lo = 2 * c->nkx0;
  r.s_lo = -1.0;
  r.s_hi = -2.0;
  r.beta = 1.0;
  HIPCHK(oa::rotate_f64(r, c->stream));
  return OA_OK;
}

// SD update with the kx=0 redundancy sync folded in (sd_twin)
oa_status st_sd(oa_ctx* c) {
  HIPCHK(oa::sd_update(c->psi, c->force, c->dtm, c->sd_twin, c->nst, c->npts, c->ld, c->fpart,
                       c->nfpart, c->stream));
  HIPCHK(oa::reduce_to(c->fpart, c->nfpart, 1.0, &c->res->fnorm2, c->stream));
  return OA_OK;
}

oa_status st_overlap(oa_ctx* c) {
  // S = 2 sum (kx0/sqrt2)^2 ... = sum_p w_p Re conj(psi_i) psi_j, w = 1 on kx=0, 2 elsewhere
  TRY(st_gram(c, c->psi, c->psi, true, 1.0, 2.0, c->S, split_on(c, kSplitS)));
  c->have_S = true;
  return OA_OK;
}

oa_status st_ns(oa_ctx* c, double tol, int max_iter) {
  if (max_iter > c->opt.ns_max_iter)
    return fail(c, OA_EINVAL, "max_iter exceeds the context's ns_max_iter");
  oa::NsArgs a;
  a.S = c->S;
  a.X0 = c->nsX0;
  a.X1 = c->nsX1;
  a.T0 = c->nsT0;
  a.T1 = c->nsT1;
  a.C = c->nsC;
  a.Tout = c->T;
  a.errpart = c->errpart;
  a.res = c->res;
  a.n = c->nst;
  a.max_iter = max_iter;
  a.tol = tol;
  a.coop = c->ns_kind;
  if (c->fp32 && c->opt.ns_fp32)
    HIPCHK(oa::ns_run_mixed(a, c->ns32, c->stream));
  else
    HIPCHK(oa::ns_run(a, c->stream));
  c->have_T = true;
  return OA_OK;
}

oa_status st_rot_psi(oa_ctx* c) {
  const bool sp = split_on(c, kSplitRot) && split_rotate_ok(c);
  TRY(st_rotate(c, c->psi, c->T, c->psi2, 0.0, 1.0, 1.0, false, false, sp, sp));
  std::swap(c->psi, c->psi2);
  return OA_OK;
}

// Forget the captured step graphs (the next graph step captures anew).
void drop_graphs(oa_ctx* c) {
  for (int p = 0; p < 2; p++) {
    if (c->gpre[p]) (void)hipGraphExecDestroy(c->gpre[p]);
    if (c->gpost[p]) (void)hipGraphExecDestroy(c->gpost[p]);
    c->gpre[p] = c->gpost[p] = nullptr;
  }
}

oa_status fetch_result(oa_ctx* c, oa_step_out* out) {
  HIPCHK(hipMemcpyAsync(c->res_host, c->res, sizeof(StepDev), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (out) {
    const StepDev& r = *c->res_host;
    out->eke = r.eke;
    out->eloc = r.eloc;
    out->ehart = c->opt.vks_from_rho ? r.ehart : 0.0;
    out->exc = c->opt.vks_from_rho ? r.exc : 0.0;
    out->enl = c->nl_natm > 0 ? r.enl : 0.0;
    out->eelec = r.eke + r.eloc + out->ehart + out->exc + out->enl;
    out->fnorm = std::sqrt(r.fnorm2);
    out->ns_iters = (r.ns_iter < 0 ? -r.ns_iter : r.ns_iter) +
                    (c->fp32 && c->opt.ns_fp32 ? r.ns32_iter : 0);
    out->ns_err = r.ns_err;
    out->resid_max = r.resid_max;
    out->resid_rms = r.resid_rms;
    out->ms = 0.0;
    if (c->timing) {
      float ms = 0.f;
      const size_t b = (size_t)c->ev_cur * OA_NTIMINGS;
      if (hipEventElapsedTime(&ms, c->ev[b], c->ev[b + 12]) == hipSuccess) out->ms = ms;
    }
  }
  return OA_OK;
}

// Mean device ms of fn() over three calls after one warm-up call, on the
// context's stream; < 0 if a call fails.
template <typename F>
double time_stage(oa_ctx* c, hipEvent_t e0, hipEvent_t e1, F&& fn) {
  if (fn() != OA_OK) return -1.0;
  if (hipEventRecord(e0, c->stream) != hipSuccess) return -1.0;
  for (int r = 0; r < 3; r++)
    if (fn() != OA_OK) return -1.0;
  if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
    return -1.0;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.0;
  return ms / 3.0;
}

// Try each candidate value of one knob (set(v) applies it), keep the
// fastest, log the times; returns the best time (ms), or -1 when there was
// no choice to make.  The data in the buffers is whatever they hold (zeros
// at creation): the kernels' time does not depend on it.
template <typename Set, typename Run>
double tune_knob(oa_ctx* c, hipEvent_t e0, hipEvent_t e1, const char* name,
                 const std::vector<int>& cand, int current, Set&& set, Run&& run) {
  if (cand.size() < 2) return -1.0;
  int best = current;
  double tbest = 1e300;
  std::string line = std::string(name) + ":";
  for (int v : cand) {
    set(v);
    const double t = time_stage(c, e0, e1, run);
    char buf[64];
    std::snprintf(buf, sizeof(buf), " %d=%.4fms", v, t);
    line += buf;
    if (t >= 0.0 && t < tbest) {
      tbest = t;
      best = v;
    }
  }
  set(best);
  line += " -> " + std::to_string(best) + (best == current ? " (heuristic)\n" : "\n");
  c->tune_log += line;
  return tbest;
}

// The measured choice of launch shapes (opt.autotune).
oa_status autotune(oa_ctx* c) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(c, OA_EHIP, "hipEventCreate failed");
  }
  FftCtx& f = c->fc;
  if (c->engine == kEngPlane) {
    HIPCHK(oa::fft2_launch(oa::FFT_ZFWD, c->nz, c->fp32, fctx(c), c->stream));
    tune_knob(
        c, e0, e1, "G->R state chunks", c->tune_nchunk, f.nchunk,
        [&](int v) {
          f.cs = (c->nst + v - 1) / v;
          f.nchunk = (c->nst + f.cs - 1) / f.cs;
        },
        [&]() -> oa_status {
          HIPCHK(oa::plane3_launch(0, c->nx, c->fp32, fctx(c), c->stream));
          HIPCHK(oa::rho_reduce(fctx(c), c->dpart, c->ndpart, c->stream, !c->opt.vks_from_rho));
          return OA_OK;
        });
    tune_knob(
        c, e0, e1, c->fuse_rho ? "R->G (+rho) state chunks" : "R->G state chunks",
        c->tune_nchunk_r, f.nchunk_r,
        [&](int v) {
          f.cs_r = (c->nst + v - 1) / v;
          f.nchunk_r = (c->nst + f.cs_r - 1) / f.cs_r;
        },
        [&]() -> oa_status {
          HIPCHK(oa::plane3_launch(c->fuse_rho ? 3 : 1, c->nx, c->fp32, fctx(c), c->stream));
          if (c->fuse_rho) {
            FftCtx fr = fctx(c);
            fr.nchunk = fr.nchunk_r;
            HIPCHK(oa::rho_reduce(fr, c->dpart, c->ndpart, c->stream));
          }
          return OA_OK;
        });
    if (c->fuse_rho) {
      // the fused pass against the separate G->R and R->G passes, each at
      // its tuned chunking (the separate R->G pass at the fused one's)
      const double t_fused = time_stage(c, e0, e1, [&]() -> oa_status {
        FftCtx fr = fctx(c);
        HIPCHK(oa::plane3_launch(3, c->nx, c->fp32, fr, c->stream));
        fr.nchunk = fr.nchunk_r;
        return oa::rho_reduce(fr, c->dpart, c->ndpart, c->stream) == hipSuccess ? OA_OK
                                                                             : OA_EHIP;
      });
      const double t_sep = time_stage(c, e0, e1, [&]() -> oa_status {
        HIPCHK(oa::plane3_launch(0, c->nx, c->fp32, fctx(c), c->stream));
        HIPCHK(oa::rho_reduce(fctx(c), c->dpart, c->ndpart, c->stream, !c->opt.vks_from_rho));
        HIPCHK(oa::plane3_launch(1, c->nx, c->fp32, fctx(c), c->stream));
        return OA_OK;
      });
      char buf[128];
      std::snprintf(buf, sizeof(buf),
                    "rho in the R->G pass (1) or a G->R pass (0): 1=%.4fms 0=%.4fms -> %d "
                    "(heuristic 1)\n",
                    t_fused, t_sep, (t_fused >= 0.0 && t_sep >= 0.0 && t_sep < t_fused) ? 0 : 1);
      c->tune_log += buf;
      if (t_fused >= 0.0 && t_sep >= 0.0 && t_sep < t_fused) c->fuse_rho = false;
    }
  }
  tune_knob(
      c, e0, e1, "S K slices", c->tune_nslice_sym, c->nslice_sym,
      [&](int v) { c->nslice_sym = v; },
      [&]() {
        return st_gram(c, c->psi, c->psi, true, 1.0 / std::sqrt(2.0), 1.0, c->S,
                       split_on(c, kSplitS));
      });
  if (c->opt.vks_from_rho && c->tune_vks.size() > 1) {
    // V_KS from the rho of the trial passes above; the energy partials are
    // summed in workgroup order, so the choice moves E_H and E_xc by rounding
    oa::VksArgs& va = c->va;
    tune_knob(
        c, e0, e1, "V_KS row pairs per x-pass workgroup", c->tune_vks, va.pr,
        [&](int v) { va.pr = v; }, [&]() { return st_vks(c); });
    tune_knob(
        c, e0, e1, "V_KS columns per y-pass workgroup", c->tune_vks, va.cb,
        [&](int v) { va.cb = v; }, [&]() { return st_vks(c); });
    tune_knob(
        c, e0, e1, "V_KS z lines per workgroup", c->tune_vks, va.lb,
        [&](int v) { va.lb = v; }, [&]() { return st_vks(c); });
  }
  tune_knob(
      c, e0, e1, "Lambda K slices", c->tune_nslice_full, c->nslice_full,
      [&](int v) { c->nslice_full = v; },
      [&]() {
        return st_gram(c, c->psi, c->force, false, 0.5, 1.0, c->lam, split_on(c, kSplitLambda));
      });
  if (c->ns_kinds.size() > 1) {
    // S = 2I converges in one iteration: the cost of one iteration either way
    std::vector<double> two((size_t)c->nst * c->nst, 0.0);
    for (int i = 0; i < c->nst; i++) two[(size_t)i * c->nst + i] = 2.0;
    HIPCHK(hipMemcpy(c->S, two.data(), two.size() * sizeof(double), hipMemcpyHostToDevice));
    tune_knob(
        c, e0, e1,
        "Newton-Schulz row blocks split over 4, 2, 1 workgroups (4, 3, 2), cooperative tiles (1), "
        "one workgroup (5) or launches (0)",
        c->ns_kinds,
        c->ns_kind, [&](int v) { c->ns_kind = v; },
        [&]() { return st_ns(c, c->opt.ns_tol, c->opt.ns_max_iter); });
  }
  if (c->nst <= 128 && c->opt.use_lambda && !c->opt.use_cg) {
    // F = 0 and psi = 0 at creation: the SD update leaves psi as it is
    tune_knob(
        c, e0, e1, "SD update in the Lambda rotate (1) or a pass of its own (0)", {1, 0}, 0,
        [&](int v) { c->sd_fused = v != 0; },
        [&]() -> oa_status {
          if (c->sd_fused) {
            TRY(st_lambda(c, true));
            return st_sd_finish(c);
          }
          TRY(st_lambda(c));
          return st_sd(c);
        });
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHK(hipStreamSynchronize(c->stream));
  // leave no state from the trial runs behind
  HIPCHK(oa::fill_zero(c->res, sizeof(StepDev), c->stream));
  if (c->opt.vks_from_rho)  // V_KS of the trials' (zero) density
    HIPCHK(oa::fill_zero(c->vks, (size_t)c->nx * c->ny * c->nz * sizeof(double), c->stream));
  c->have_S = c->have_T = false;
  HIPCHK(hipStreamSynchronize(c->stream));
  return OA_OK;
}

}  // namespace

extern "C" {

void oa_default_options(oa_options* o) {
  o->precision = OA_FP64;
  o->ns_max_iter = 30;
  o->ns_tol = 1e-15;
  o->dt = 5.0;
  o->tau_mass = 25.0;
  o->ecut_mass = 2.0;
  o->use_lambda = 1;
  o->use_cg = 0;
  o->want_resid = 1;
  o->fft_mode = 0;
  o->use_graph = 0;
  o->ns_mode = 0;
  o->vks_from_rho = 0;
  o->autotune = 0;
  o->gemm_split = 1;
  o->ns_fp32 = 0;
  o->fft_full = 0;
  o->psir_keep = 0;
}

int oa_kvecs(double ecut_ry, const double hmat[9], int* kx, int* ky, int* kz) {
  std::vector<int> a, b, cc;
  enumerate_kvecs(ecut_ry, hmat, a, b, cc);
  if (kx && ky && kz) {
    std::copy(a.begin(), a.end(), kx);
    std::copy(b.begin(), b.end(), ky);
    std::copy(cc.begin(), cc.end(), kz);
  }
  return (int)a.size();
}

}  // extern "C"

// dry: lay the context out and count its device bytes without allocating,
// uploading or launching anything (oa_ctx_bytes)
static oa_status create_impl(int device, const oa_system* sys, const oa_options* opt, bool dry,
                             oa_ctx** out) {
  if (!sys || !out) return OA_EINVAL;
  *out = nullptr;
  oa_ctx* c = new oa_ctx();
  c->dry = dry;
  auto bail = [&](oa_status s) {
    if (!dry) std::fprintf(stderr, "oa_create: %s\n", c->err.c_str());
    oa_destroy(c);
    return s;
  };
  c->device = device;
  if (opt) c->opt = *opt;
  else oa_default_options(&c->opt);
  if (c->opt.ns_max_iter < 1) c->opt.ns_max_iter = 1;
  c->fp32 = (c->opt.precision == OA_MIXED);
  c->nx = sys->grid[0];
  c->ny = sys->grid[1];
  c->nz = sys->grid[2];
  c->nst = sys->nstates;
  c->ecut_ry = sys->ecut_ry;
  std::memcpy(c->hmat, sys->hmat, sizeof(c->hmat));
  c->vol = std::fabs(inv3(c->hmat, c->hmati));
  if (c->nx < 2 || c->ny < 2 || c->nz < 2 || c->nst < 1) {
    c->err = "bad grid or state count";
    return bail(OA_EINVAL);
  }
  if (c->opt.use_cg && !c->opt.use_lambda) {
    // computeCgOverlap needs the Lambda-corrected forces (CP_State_GSpacePlane.C:2349-2353)
    c->err = "conjugate gradient needs use_lambda";
    return bail(OA_EINVAL);
  }
  if (sys->npts > 0) {
    if (!sys->kx || !sys->ky || !sys->kz) {
      c->err = "npts > 0 needs kx/ky/kz";
      return bail(OA_EINVAL);
    }
    c->kx.assign(sys->kx, sys->kx + sys->npts);
    c->ky.assign(sys->ky, sys->ky + sys->npts);
    c->kz.assign(sys->kz, sys->kz + sys->npts);
  } else {
    enumerate_kvecs(c->ecut_ry, c->hmat, c->kx, c->ky, c->kz);
  }
  c->npts = (int)c->kx.size();
  c->ld = ((c->npts + 31) / 32) * 32;
  if (hipSetDevice(device) != hipSuccess) {
    c->err = "hipSetDevice failed";
    return bail(OA_EHIP);
  }
  if (!dry && hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    c->err = "stream creation failed";
    return bail(OA_EHIP);
  }
  if (!dry) {
    // the arena: the device bytes a sizing run of the same system counts
    oa_ctx* d = nullptr;
    if (create_impl(device, sys, opt, true, &d) == OA_OK && d) {
      if (hipMalloc((void**)&c->arena, d->bytes) == hipSuccess) {
        c->arena_size = d->bytes;
      } else {  // separate blocks instead
        (void)hipGetLastError();
        c->arena = nullptr;
      }
    }
    oa_destroy(d);
  }
  oa_status s = build_layout(c);
  if (s != OA_OK) return bail(s);
  s = alloc_state(c, sys->occ);
  if (s != OA_OK) return bail(s);
  const size_t need = oa::fft_max_smem(c->fc, c->fp32);
  if (need > 64 * 1024) {
    c->err = "FFT workgroup needs " + std::to_string(need) + " bytes of LDS";
    return bail(OA_EINVAL);
  }
  if (c->opt.autotune && !dry) {
    s = autotune(c);
    if (s != OA_OK) return bail(s);
  }
  *out = c;
  return OA_OK;
}

extern "C" {

oa_status oa_create(int device, const oa_system* sys, const oa_options* opt, oa_ctx** out) {
  return create_impl(device, sys, opt, false, out);
}

size_t oa_ctx_bytes(const oa_system* sys, const oa_options* opt) {
  oa_ctx* c = nullptr;
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (create_impl(dev, sys, opt, true, &c) != OA_OK || !c) return 0;
  const size_t b = c->bytes;
  oa_destroy(c);
  return b;
}

size_t oa_ctx_device_bytes(const oa_ctx* c) { return c ? c->bytes : 0; }

}  // extern "C"

hipStream_t oa_ctx_stream(oa_ctx* c) { return c ? c->stream : nullptr; }
int oa_ctx_nstates(const oa_ctx* c) { return c ? c->nst : 0; }
void oa_ctx_grid(const oa_ctx* c, int grid[3]) {
  grid[0] = c ? c->nx : 0;
  grid[1] = c ? c->ny : 0;
  grid[2] = c ? c->nz : 0;
}

extern "C" {

oa_status oa_destroy(oa_ctx* c) {
  if (!c) return OA_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graphs(c);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->arena) (void)hipFree(c->arena);
  if (c->stage) (void)hipFree(c->stage);
  if (c->res_host) (void)hipHostFree(c->res_host);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  c->ev.clear();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return OA_OK;
}

const char* oa_last_error(const oa_ctx* c) { return c ? c->err.c_str() : "null context"; }

const char* oa_tuning_log(const oa_ctx* c) { return c ? c->tune_log.c_str() : ""; }

oa_status oa_layout_info(const oa_ctx* c, int info[4]) {
  if (!c || !info) return OA_EINVAL;
  info[0] = c->npts;
  info[1] = c->nlines;
  info[2] = c->nplane_x;
  info[3] = c->nkx0;
  return OA_OK;
}

oa_status oa_get_config(const oa_ctx* c, int cfg[8]) {
  if (!c || !cfg) return OA_EINVAL;
  const FftCtx& f = c->fc;
  const bool fused = c->engine == kEngFused, plane = c->engine == kEngPlane;
  const bool fft2 = c->engine == kEngFft2 || plane;
  cfg[0] = c->fuse_rho ? 4 : c->engine;
  cfg[1] = fft2 ? oa::fft2_ntr(c->nz) : f.lpb;
  cfg[2] = fft2 ? oa::fft2_ntr(c->ny) : f.kc;
  cfg[3] = fft2 ? oa::fft2_ntr(c->nx) : f.npair;
  cfg[4] = fused ? f.kcp : plane ? f.cs_r : 0;
  cfg[5] = fused ? f.pcp : plane ? f.nchunk_r : 0;
  cfg[6] = (fused || fft2) ? f.cs : 0;
  cfg[7] = (fused || fft2) ? f.nchunk : 0;
  return OA_OK;
}

oa_status oa_get_kvecs(const oa_ctx* c, int* kx, int* ky, int* kz) {
  if (!c || !kx || !ky || !kz) return OA_EINVAL;
  std::copy(c->kx.begin(), c->kx.end(), kx);
  std::copy(c->ky.begin(), c->ky.end(), ky);
  std::copy(c->kz.begin(), c->kz.end(), kz);
  return OA_OK;
}

static oa_status copy_g(oa_ctx* c, double2* dev, const double* host, bool in) {
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpy2DAsync(in ? (void*)dev : (void*)host, in ? c->ld * sizeof(double2) : c->npts * sizeof(double2),
                          in ? (const void*)host : (const void*)dev,
                          in ? c->npts * sizeof(double2) : c->ld * sizeof(double2),
                          c->npts * sizeof(double2), c->nst,
                          in ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return OA_OK;
}

oa_status oa_psi_set(oa_ctx* c, const double* psi) {
  if (!c || !psi) return OA_EINVAL;
  if (c->opt.use_cg) {  // a new minimisation: zero direction, iteration 0
    HIPCHK(hipSetDevice(c->device));
    const size_t gsz = (size_t)c->nst * c->ld * sizeof(double2);
    HIPCHK(oa::fill_zero(c->cgdir, gsz, c->stream));
    HIPCHK(oa::fill_zero(c->cgs, sizeof(oa::CgDev), c->stream));
  }
  return copy_g(c, c->psi, psi, true);
}
oa_status oa_psi_get(oa_ctx* c, double* psi) {
  if (!c || !psi) return OA_EINVAL;
  return copy_g(c, c->psi, psi, false);
}
oa_status oa_force_get(oa_ctx* c, double* f) {
  if (!c || !f) return OA_EINVAL;
  return copy_g(c, c->force, f, false);
}
oa_status oa_force_set(oa_ctx* c, const double* f) {
  if (!c || !f) return OA_EINVAL;
  return copy_g(c, c->force, f, true);
}

oa_status oa_vks_set(oa_ctx* c, const double* v) {
  if (!c || !v) return OA_EINVAL;
  if (c->opt.vks_from_rho)
    return fail(c, OA_ESTATE, "this context builds V from rho (vks_from_rho)");
  HIPCHK(hipSetDevice(c->device));
  TRY(upload(c, c->vks, v, (size_t)c->nx * c->ny * c->nz * sizeof(double)));
  c->have_v = true;
  return OA_OK;
}

oa_status oa_local_pseudo_set(oa_ctx* c, const oa_local_pseudo* lp) {
  if (!c || !lp) return OA_EINVAL;
  if (!c->opt.vks_from_rho)
    return fail(c, OA_ESTATE, "oa_local_pseudo_set needs a vks_from_rho context");
  if (lp->natoms < 1 || lp->nspecies < 1 || lp->nspline < 1 || !(lp->dg > 0.0) || !lp->pos ||
      !lp->species || !lp->vps || !lp->vps_g0)
    return fail(c, OA_EINVAL, "bad local pseudopotential description");
  for (int i = 0; i < lp->natoms; i++)
    if (lp->species[i] < 0 || lp->species[i] >= lp->nspecies)
      return fail(c, OA_EINVAL, "atom species out of range");
  // the density sphere's largest |G| must lie on the spline grid
  if (lp->gmin + (lp->nspline - 1) * lp->dg < std::sqrt(c->va.ecut4))
    return fail(c, OA_EINVAL, "spline table ends below the density cutoff");
  HIPCHK(hipSetDevice(c->device));
  const size_t bpos = sizeof(double) * 3 * lp->natoms, bsp = sizeof(int) * lp->natoms;
  const size_t bvps = sizeof(double) * 4 * (size_t)lp->nspecies * lp->nspline;
  const size_t bg0 = sizeof(double) * lp->nspecies;
  char* tmp = nullptr;
  const size_t off_sp = (bpos + 255) & ~(size_t)255;
  const size_t off_vps = off_sp + ((bsp + 255) & ~(size_t)255);
  const size_t off_g0 = off_vps + ((bvps + 255) & ~(size_t)255);
  HIPCHK(hipStreamSynchronize(c->stream));  // the staging block may be in use
  TRY(stage_buf(c, off_g0 + bg0, &tmp));
  oa::PseudoDev p{};
  p.natoms = lp->natoms;
  p.nspecies = lp->nspecies;
  p.nspline = lp->nspline;
  p.gmin = lp->gmin;
  p.dg = lp->dg;
  p.pos = reinterpret_cast<const double*>(tmp);
  p.species = reinterpret_cast<const int*>(tmp + off_sp);
  p.vps = reinterpret_cast<const double*>(tmp + off_vps);
  p.vps_g0 = reinterpret_cast<const double*>(tmp + off_g0);
  hipError_t e = hipMemcpy(tmp, lp->pos, bpos, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(tmp + off_sp, lp->species, bsp, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(tmp + off_vps, lp->vps, bvps, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(tmp + off_g0, lp->vps_g0, bg0, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = oa::vloc_build(c->va, p, c->vloc, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  HIPCHK(e);
  return OA_OK;
}

oa_status oa_vks_build(oa_ctx* c) {
  if (!c) return OA_EINVAL;
  if (!c->opt.vks_from_rho)
    return fail(c, OA_ESTATE, "oa_vks_build needs a vks_from_rho context");
  HIPCHK(hipSetDevice(c->device));
  TRY(st_vks(c));
  HIPCHK(hipStreamSynchronize(c->stream));
  return OA_OK;
}

oa_status oa_vks_get(oa_ctx* c, double* v) {
  if (!c || !v) return OA_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(v, c->vks, (size_t)c->nx * c->ny * c->nz * sizeof(double),
                   hipMemcpyDeviceToHost));
  return OA_OK;
}

oa_status oa_nonlocal_set(oa_ctx* c, const oa_nonlocal* nl) {
  if (!c || !nl) return OA_EINVAL;
  if (nl->natoms < 1 || nl->nspecies < 1 || nl->nspline < 1 || !(nl->dg > 0.0) || !nl->pos ||
      !nl->species || !nl->proj || !nl->proj_g0 || !nl->vnorm)
    return fail(c, OA_EINVAL, "bad non-local pseudopotential description");
  for (int i = 0; i < nl->natoms; i++)
    if (nl->species[i] < 0 || nl->species[i] >= nl->nspecies)
      return fail(c, OA_EINVAL, "non-local atom species out of range");
  // the wavefunction sphere's largest |G| must lie on the spline grid
  if (nl->gmin + (nl->nspline - 1) * nl->dg < std::sqrt(c->ecut_ry))
    return fail(c, OA_EINVAL, "projector table ends below the wavefunction cutoff");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  const int na = nl->natoms;
  if (na > c->nl_cap) {  // buffers for this many atoms (kept until destroy)
    const int ldr = 2 * c->ld;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) !=
            hipSuccess || ncu <= 0)
      ncu = 256;
    const int ns = oa::gram_slices(std::max(c->nst, 4 * na), 2 * c->npts, 2 * c->nkx0, 0, ncu);
    TRY(dalloc(c, &c->nl_sf, (size_t)4 * na * ldr * sizeof(double)));
    TRY(dalloc(c, &c->nl_z, (size_t)c->nst * 4 * na * sizeof(double)));
    TRY(dalloc(c, &c->nl_mt, (size_t)na * c->nst * sizeof(double)));
    TRY(dalloc(c, &c->nl_vn, (size_t)na * sizeof(double)));
    TRY(dalloc(c, &c->nl_part, (size_t)ns * c->nst * 4 * na * sizeof(double)));
    TRY(dalloc(c, &c->nl_fatm, (size_t)na * 3 * sizeof(double)));
    c->nl_cap = na;
    c->nl_nslice = ns;
  }
  // the table of this atom set, rows [4 na][2 ld] (padding points zero)
  const size_t ldr = 2 * (size_t)c->ld;
  HIPCHK(oa::fill_zero(c->nl_sf, (size_t)4 * na * ldr * sizeof(double), c->stream));
  const size_t bk = sizeof(int) * c->npts, bpos = sizeof(double) * 3 * na;
  const size_t bsp = sizeof(int) * na;
  const size_t bproj = sizeof(double) * 4 * (size_t)nl->nspecies * nl->nspline;
  const size_t bg0 = sizeof(double) * nl->nspecies;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_ky = up(bk), o_kz = o_ky + up(bk), o_pos = o_kz + up(bk);
  const size_t o_sp = o_pos + up(bpos), o_proj = o_sp + up(bsp), o_g0 = o_proj + up(bproj);
  char* tmp = nullptr;
  TRY(stage_buf(c, o_g0 + bg0, &tmp));
  oa::NlDev d{};
  d.natoms = na;
  d.nspecies = nl->nspecies;
  d.nspline = nl->nspline;
  d.npts = c->npts;
  d.ldr = (int)ldr;
  d.gmin = nl->gmin;
  d.dg = nl->dg;
  d.y00 = 1.0 / std::sqrt(4.0 * M_PI);
  std::copy(c->hmati, c->hmati + 9, d.hmati);
  d.kx = reinterpret_cast<const int*>(tmp);
  d.ky = reinterpret_cast<const int*>(tmp + o_ky);
  d.kz = reinterpret_cast<const int*>(tmp + o_kz);
  d.pos = reinterpret_cast<const double*>(tmp + o_pos);
  d.species = reinterpret_cast<const int*>(tmp + o_sp);
  d.proj = reinterpret_cast<const double*>(tmp + o_proj);
  d.proj_g0 = reinterpret_cast<const double*>(tmp + o_g0);
  std::vector<double> vn(na);
  for (int i = 0; i < na; i++) vn[i] = nl->vnorm[nl->species[i]] / c->vol;
  hipError_t e = hipMemcpy(tmp, c->kx.data(), bk, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(tmp + o_ky, c->ky.data(), bk, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(tmp + o_kz, c->kz.data(), bk, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(tmp + o_pos, nl->pos, bpos, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(tmp + o_sp, nl->species, bsp, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(tmp + o_proj, nl->proj, bproj, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(tmp + o_g0, nl->proj_g0, bg0, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(c->nl_vn, vn.data(), sizeof(double) * na, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = oa::nl_sf_build(d, c->nl_sf, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  HIPCHK(e);
  c->nl_natm = na;
  drop_graphs(c);  // captured steps lack (or hold the old) non-local stage
  return OA_OK;
}

oa_status oa_nonlocal_get(oa_ctx* c, double* z, double* enl, double* fatoms) {
  if (!c) return OA_EINVAL;
  if (c->nl_natm < 1) return fail(c, OA_ESTATE, "no non-local pseudopotential set");
  HIPCHK(hipSetDevice(c->device));
  const int na = c->nl_natm, nb = fatoms ? 4 * na : na;
  TRY(st_nl_z(c, fatoms != nullptr));
  HIPCHK(oa::nl_energy(c->nl_z, nb, c->nst, na, c->nl_vn, c->occ, c->nl_mt, &c->res->enl,
                       c->stream));
  if (fatoms) HIPCHK(oa::nl_atom_forces(c->nl_z, c->nst, na, c->nl_vn, c->nl_fatm, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (z) {
    std::vector<double> h((size_t)c->nst * nb);
    HIPCHK(hipMemcpy(h.data(), c->nl_z, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int s = 0; s < c->nst; s++)
      std::copy(h.begin() + (size_t)s * nb, h.begin() + (size_t)s * nb + na, z + (size_t)s * na);
  }
  if (fatoms)
    HIPCHK(hipMemcpy(fatoms, c->nl_fatm, sizeof(double) * 3 * na, hipMemcpyDeviceToHost));
  TRY(fetch_result(c, nullptr));
  if (enl) *enl = c->res_host->enl;
  return OA_OK;
}

oa_status oa_nonlocal_force(oa_ctx* c) {
  if (!c) return OA_EINVAL;
  if (c->nl_natm < 1) return fail(c, OA_ESTATE, "no non-local pseudopotential set");
  HIPCHK(hipSetDevice(c->device));
  TRY(st_nonlocal(c));
  HIPCHK(hipStreamSynchronize(c->stream));
  return OA_OK;
}

oa_status oa_fft_g2r(oa_ctx* c) {
  if (!c) return OA_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  TRY(st_g2r(c));
  HIPCHK(hipStreamSynchronize(c->stream));
  return OA_OK;
}

oa_status oa_fft_g2r_states(oa_ctx* c, int s0, int ns, int first) {
  if (!c) return OA_EINVAL;
  TRY(check_range(c, s0, ns));
  HIPCHK(hipSetDevice(c->device));
  FftCtx f = range_ctx(c, s0, ns);
  HIPCHK(oa::fft2_launch(oa::FFT_ZFWD, c->nz, c->fp32, f, c->stream));
  HIPCHK(oa::plane3_launch(0, c->nx, c->fp32, f, c->stream));
  HIPCHK(oa::rho_reduce(f, c->dpart, c->ndpart, c->stream, false, first == 0));
  HIPCHK(hipStreamSynchronize(c->stream));
  return OA_OK;
}

oa_status oa_psir_get(oa_ctx* c, int s, double* out) {
  if (!c || !out || s < 0 || s >= c->nst) return OA_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  const size_t N = (size_t)c->nx * c->ny * c->nz;
  size_t off = (size_t)s * N;
  if (c->engine == kEngPlane) {
    // psi(r) is not kept: recompute state s from the z-pass output
    FftCtx f = fctx(c);
    f.psir_state = s;
    HIPCHK(oa::plane3_launch(2, c->nx, c->fp32, f, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    off = 0;
  }
  if (!c->fp32) {
    HIPCHK(hipMemcpy(out, (double*)c->psir + off, N * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::vector<float> tmp(N);
    HIPCHK(hipMemcpy(tmp.data(), (float*)c->psir + off, N * sizeof(float),
                     hipMemcpyDeviceToHost));
    for (size_t i = 0; i < N; i++) out[i] = tmp[i];
  }
  return OA_OK;
}

oa_status oa_density_get(oa_ctx* c, double* rho) {
  if (!c || !rho) return OA_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(rho, c->rho, (size_t)c->nx * c->ny * c->nz * sizeof(double),
                   hipMemcpyDeviceToHost));
  return OA_OK;
}

oa_status oa_vpsi_r2g(oa_ctx* c, double* eke) {
  if (!c) return OA_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  TRY(st_r2g(c));
  TRY(fetch_result(c, nullptr));
  if (eke) *eke = c->res_host->eke;
  return OA_OK;
}

oa_status oa_vpsi_r2g_states(oa_ctx* c, int s0, int ns, double* eke) {
  if (!c) return OA_EINVAL;
  TRY(check_range(c, s0, ns));
  HIPCHK(hipSetDevice(c->device));
  FftCtx f = range_ctx(c, s0, ns);
  HIPCHK(oa::plane3_launch(1, c->nx, c->fp32, f, c->stream));
  f.zbuf = f.zout;  // the z pass back reads the R->G plane output
  HIPCHK(oa::fft2_launch(oa::FFT_ZBWD, c->nz, c->fp32, f, c->stream));
  HIPCHK(oa::reduce_to(f.eke_part, ns * c->nzblk, 1.0, &c->res->eke, c->stream));
  TRY(fetch_result(c, nullptr));
  if (eke) *eke = c->res_host->eke;
  return OA_OK;
}

oa_status oa_lambda(oa_ctx* c, double* lambda) {
  if (!c) return OA_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  TRY(st_lambda(c));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (lambda)
    HIPCHK(hipMemcpy(lambda, c->lam, (size_t)c->nst * c->nst * sizeof(double),
                     hipMemcpyDeviceToHost));
  return OA_OK;
}

oa_status oa_sd_update(oa_ctx* c, double* fnorm) {
  if (!c) return OA_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  TRY(st_sd(c));
  TRY(fetch_result(c, nullptr));
  if (fnorm) *fnorm = std::sqrt(c->res_host->fnorm2);
  return OA_OK;
}

oa_status oa_overlap(oa_ctx* c, double* S) {
  if (!c) return OA_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  TRY(st_overlap(c));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (S)
    HIPCHK(hipMemcpy(S, c->S, (size_t)c->nst * c->nst * sizeof(double), hipMemcpyDeviceToHost));
  return OA_OK;
}

oa_status oa_ortho_ns(oa_ctx* c, double tol, int max_iter, int* iters, double* err, double* T) {
  if (!c || max_iter < 1) return OA_EINVAL;
  if (!c->have_S) return fail(c, OA_ESTATE, "oa_ortho_ns before oa_overlap");
  HIPCHK(hipSetDevice(c->device));
  TRY(st_ns(c, tol, max_iter));
  TRY(fetch_result(c, nullptr));
  const int it = c->res_host->ns_iter;
  if (iters) *iters = it < 0 ? -it : it;
  if (err) *err = c->res_host->ns_err;
  if (T)
    HIPCHK(hipMemcpy(T, c->T, (size_t)c->nst * c->nst * sizeof(double), hipMemcpyDeviceToHost));
  if (it < 0) return fail(c, OA_ENOCONV, "invsqr_tolerance not met within invsqr_max_iter");
  return OA_OK;
}

oa_status oa_tmat_set(oa_ctx* c, const double* T) {
  if (!c || !T) return OA_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  TRY(upload(c, c->T, T, (size_t)c->nst * c->nst * sizeof(double)));
  c->have_T = true;
  return OA_OK;
}

oa_status oa_rotate(oa_ctx* c) {
  if (!c) return OA_EINVAL;
  if (!c->have_T) return fail(c, OA_ESTATE, "oa_rotate before a T is available");
  HIPCHK(hipSetDevice(c->device));
  TRY(st_rot_psi(c));
  HIPCHK(hipStreamSynchronize(c->stream));
  return OA_OK;
}

}  // extern "C"

namespace {

// The step before Newton-Schulz (GSpaceDriver order, gspace.ci:106-190):
// G->R, rho, V psi, R->G, EKE and kinetic force, [Lambda], SD update,
// redundancy sync, S.  marks: record the stage events.
oa_status enqueue_pre(oa_ctx* c, bool marks) {
  auto mk = [&](int i) {
    if (marks) mark(c, i);
  };
  if (c->fuse_rho) {
    TRY(st_g2r2g_rho(c, marks ? 1 : -1));  // marks 1..6, then the EKE / E_loc reductions
  } else {
    TRY(st_g2r(c, marks ? 1 : -1));  // marks 1..3
    TRY(st_r2g(c, marks ? 1 : -1));  // marks 4..6, then the EKE / E_loc reductions
  }
  if (c->nl_natm > 0) TRY(st_nonlocal(c));  // timed with the Lambda stage
  if (c->opt.use_cg) {
    TRY(st_lambda_cg(c));
    mk(7);
  } else if (c->opt.use_lambda && c->sd_fused) {
    TRY(st_lambda(c, true));  // Lambda projection and SD update in one pass
    mk(7);
    TRY(st_sd_finish(c));
  } else {
    if (c->opt.use_lambda) TRY(st_lambda(c));
    mk(7);
    TRY(st_sd(c));
  }
  mk(8);
  TRY(st_overlap(c));
  return OA_OK;
}

// Psi T into the other buffer (the host pointers swap), then the residual
oa_status enqueue_post(oa_ctx* c, bool marks) {
  TRY(st_rot_psi(c));
  if (marks) mark(c, 11);
  if (c->opt.want_resid) {
    TRY(st_gram(c, c->psi, c->psi, true, 1.0, 2.0, c->S2));
    HIPCHK(oa::resid_stats(c->S2, c->nst, c->res, c->stream));
  }
  return OA_OK;
}

// Capture one half of the step for the current parity into an executable
// graph.  Nothing runs; the psi/psi2 swap of the post half is undone.
oa_status capture(oa_ctx* c, bool post, hipGraphExec_t* exec) {
  double2 *p0 = c->psi, *p1 = c->psi2;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const oa_status s = post ? enqueue_post(c, false) : enqueue_pre(c, false);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  c->psi = p0;
  c->psi2 = p1;
  if (s != OA_OK) {
    if (e == hipSuccess && g) (void)hipGraphDestroy(g);
    return s;
  }
  HIPCHK(e);
  const hipError_t ei = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIPCHK(ei);
  return OA_OK;
}

}  // namespace

extern "C" {

oa_status oa_step(oa_ctx* c, oa_step_out* out) {
  if (!c) return OA_EINVAL;
  if (!c->have_v && !c->opt.vks_from_rho)
    return fail(c, OA_ESTATE, "oa_step before oa_vks_set");
  HIPCHK(hipSetDevice(c->device));
  if (c->timing) c->ev_cur = c->ev_count % oa_ctx::kRing;
  // graph mode from the second step on (the first one runs eagerly and sets
  // up every kernel attribute outside any capture)
  const bool graph = c->opt.use_graph && c->steps_run > 0;
  mark(c, 0);
  if (graph) {
    const int p = (c->psi == c->psi_buf[0]) ? 0 : 1;
    if (!c->gpre[p]) TRY(capture(c, false, &c->gpre[p]));
    if (!c->gpost[p]) TRY(capture(c, true, &c->gpost[p]));
    HIPCHK(hipGraphLaunch(c->gpre[p], c->stream));
    for (int i = 1; i <= 8; i++) mark(c, i);
    TRY(st_ns(c, c->opt.ns_tol, c->opt.ns_max_iter));
    mark(c, 9);
    HIPCHK(hipGraphLaunch(c->gpost[p], c->stream));
    std::swap(c->psi, c->psi2);  // what the captured rotate did
    c->have_S = c->have_T = true;
    mark(c, 10);
    mark(c, 11);
  } else {
    TRY(enqueue_pre(c, true));
    mark(c, 9);
    TRY(st_ns(c, c->opt.ns_tol, c->opt.ns_max_iter));
    mark(c, 10);
    TRY(enqueue_post(c, true));
  }
  mark(c, 12);
  c->steps_run++;
  if (c->timing) c->ev_count++;
  if (out) {
    TRY(fetch_result(c, out));
    if (c->res_host->ns_iter < 0)
      return fail(c, OA_ENOCONV, "invsqr_tolerance not met within invsqr_max_iter");
  }
  return OA_OK;
}

oa_status oa_residual(oa_ctx* c, double* rmax, double* rrms) {
  if (!c) return OA_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  TRY(st_gram(c, c->psi, c->psi, true, 1.0, 2.0, c->S2));
  HIPCHK(oa::resid_stats(c->S2, c->nst, c->res, c->stream));
  TRY(fetch_result(c, nullptr));
  if (rmax) *rmax = c->res_host->resid_max;
  if (rrms) *rrms = c->res_host->resid_rms;
  return OA_OK;
}

oa_status oa_sync(oa_ctx* c) {
  if (!c) return OA_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  return OA_OK;
}

oa_status oa_step_result(oa_ctx* c, oa_step_out* out) {
  if (!c || !out) return OA_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  TRY(fetch_result(c, out));
  if (c->res_host->ns_iter < 0)
    return fail(c, OA_ENOCONV, "invsqr_tolerance not met within invsqr_max_iter");
  return OA_OK;
}

oa_status oa_set_timing(oa_ctx* c, int on) {
  if (!c || on < 0 || on > 2) return OA_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  if (on) c->timing_level = on;
  if (on && c->ev.empty()) {
    c->ev.assign((size_t)oa_ctx::kRing * OA_NTIMINGS, nullptr);
    for (auto& e : c->ev) HIPCHK(hipEventCreate(&e));
  }
  c->timing = on != 0;
  c->ev_count = 0;
  c->ev_cur = 0;
  return OA_OK;
}

// Average over the last min(steps timed, ring size) steps.
oa_status oa_timings(oa_ctx* c, double* ms, int n) {
  if (!c || !ms || n < 1) return OA_EINVAL;
  if (!c->timing || c->ev_count == 0) return fail(c, OA_ESTATE, "no timed step");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  const int nsteps = std::min(c->ev_count, (int)oa_ctx::kRing);
  for (int i = 0; i < n && i < OA_NTIMINGS; i++) {
    const int a = (i == OA_NTIMINGS - 1) ? 0 : i;
    const int b = (i == OA_NTIMINGS - 1) ? 12 : i + 1;
    if (!mark_on(c, a) || !mark_on(c, b)) {  // not recorded at this level
      ms[i] = 0.0;
      continue;
    }
    double acc = 0.0;
    for (int k = 0; k < nsteps; k++) {
      const size_t base = (size_t)((c->ev_count - 1 - k) % oa_ctx::kRing) * OA_NTIMINGS;
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, c->ev[base + a], c->ev[base + b]));
      acc += t;
    }
    ms[i] = acc / nsteps;
  }
  return OA_OK;
}

}  // extern "C"
