// OaStepBackend.h -- the node-level shim a maintainer adds to the reference
// tree (src_charm_driver/gpu/OaStepBackend.h) so that the chares hand their
// state FFT, density, PairCalculator and Ortho work to liboa_step.so.  The
// methods keep the meaning of the reference calls they replace; any status
// other than OA_OK ends the run through CkAbort, the reference's own error
// policy (fftCache.C:115-120, ortho.C:129-133).  Inside Charm++ CkAbort
// comes from charm++.h; elsewhere the including file provides it.
#pragma once

#include <vector>

extern "C" {
#include "oa_step.h"
}

void CkAbort(const char *msg);

class OaStepBackend {
  oa_ctx *ctx_ = nullptr;

  static void check(oa_status s, oa_ctx *c) {
    if (s != OA_OK) CkAbort(c ? oa_last_error(c) : "oa_create failed");
  }

 public:
  // cell, grid, cutoff and the stored k-vectors in the reference order
  // (setkvec3d_sm + flip_data_set), one context per system and GPU
  OaStepBackend(const double hmat[9], const int grid[3], double ecut_ry, int nstates,
                const std::vector<int> &kx, const std::vector<int> &ky,
                const std::vector<int> &kz, double dt, bool mixed, int device = 0) {
    oa_system sys{};
    for (int i = 0; i < 9; i++) sys.hmat[i] = hmat[i];
    for (int i = 0; i < 3; i++) sys.grid[i] = grid[i];
    sys.ecut_ry = ecut_ry;
    sys.nstates = nstates;
    sys.occ = nullptr;
    sys.npts = (int)kx.size();
    sys.kx = kx.data();
    sys.ky = ky.data();
    sys.kz = kz.data();
    oa_options opt;
    oa_default_options(&opt);
    opt.dt = dt;
    opt.precision = mixed ? OA_MIXED : OA_FP64;
    check(oa_create(device, &sys, &opt, &ctx_), nullptr);
  }
  ~OaStepBackend() { oa_destroy(ctx_); }
  OaStepBackend(const OaStepBackend &) = delete;
  OaStepBackend &operator=(const OaStepBackend &) = delete;

  // GSpaceDriver (gspace.ci:106-190): one minimisation iteration of all states
  void step(const double *vks, oa_step_out *out) {
    check(oa_vks_set(ctx_, vks), ctx_);
    check(oa_step(ctx_, out), ctx_);
  }
  // stage by stage, keeping the chares' message flow
  void gtoR(double *rho) {  // doStpFFTGtoR_* + the density reduction
    check(oa_fft_g2r(ctx_), ctx_);
    check(oa_density_get(ctx_, rho), ctx_);
  }
  double rtoG(const double *vks) {  // unpackVks + doStpFFTRtoG_* + doIFFT + CP_eke_calc
    double eke = 0.0;
    check(oa_vks_set(ctx_, vks), ctx_);
    check(oa_vpsi_r2g(ctx_, &eke), ctx_);
    return eke;
  }
  void lambda(double *lam) { check(oa_lambda(ctx_, lam), ctx_); }
  double integrate() {  // CP_integrate_min_STD + the redundancy sync
    double f = 0.0;
    check(oa_sd_update(ctx_, &f), ctx_);
    return f;
  }
  void ortho(double *S, double *T) {  // multiplyForward + Ortho + bwMultiplyHelper
    int it = 0;
    double err = 0.0;
    check(oa_overlap(ctx_, S), ctx_);
    check(oa_ortho_ns(ctx_, 1e-15, 30, &it, &err, T), ctx_);
    check(oa_rotate(ctx_), ctx_);
  }
  void psiIn(const double *psi) { check(oa_psi_set(ctx_, psi), ctx_); }
  void psiOut(double *psi) { check(oa_psi_get(ctx_, psi), ctx_); }
};
