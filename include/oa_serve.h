/*
 * oa_serve.h -- request loop for independent electronic minimisations on one
 * GPU (openatom_amd/liboa_step.so).
 *
 * OpenAtom runs several independent CP instances in one job -- path-integral
 * beads, k-points, tempering replicas, spins -- each a full set of chares
 * addressed by an Uber index and scheduled together by the Charm++ runtime
 * (uber/Uber.h, UberCollection; cpaimd.C instance set-up).  This interface
 * is the one-GPU counterpart for independent fixed-ion minimisations of
 * different sizes: a native scheduler thread keeps every admitted request's
 * wavefunctions resident in HBM, admits queued requests while the HBM budget
 * and the slot count allow, advances every admitted request one CP step at a
 * time on its own HIP stream (so their kernels overlap on the device), and
 * retires a request when |F| <= f_tol (the tol_cp_min exit of
 * CP_State_GSpacePlane.C:158-191) or after max_steps.  A finished request's
 * context stays resident (within the HBM budget) for the next request of
 * the same shape: cell, grid, cutoff, states, k-vectors, occupations and
 * options; it is released when another shape needs the memory.  A request's result
 * does not depend on what else runs: each context owns its buffers and
 * stream, and every kernel is deterministic.
 *
 * Threading: submit/poll/wait/drain may be called from any host thread; the
 * contexts are owned by the scheduler thread.
 */
#ifndef OA_SERVE_H
#define OA_SERVE_H

#include <stddef.h>

#include "oa_step.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oa_server oa_server;

typedef struct {
  oa_system sys;        /* the request's cell, grid, cutoff, states (k-vectors copied) */
  oa_options opt;       /* step options (precision, dt, use_cg, ...); vks_from_rho is
                           set from vks / pseudo below */
  const double *psi0;   /* [nstates][npts] complex starting coefficients (copied) */
  const double *vks;    /* [Nz][Ny][Nx] fixed V(r) (copied), or NULL for a
                           self-consistent request (pseudo) */
  double f_tol;         /* converged when |F| <= f_tol */
  int max_steps;        /* OA_ENOCONV after this many steps */
  double *psi_out;      /* optional [nstates][npts] final coefficients; must stay valid
                           until the request's result is returned */
  long long tag;        /* caller's id, returned with the result */
  double fp64_below;    /* mixed precision only: continue in FP64, from the current
                           coefficients with a fresh CG direction, once |F| <= fp64_below or
                           |F| has not improved for 50 steps (the FP32 FFT chain sets a
                           floor on |F|); 0 disables the switch */
  const oa_local_pseudo *pseudo; /* with vks NULL: the fixed ions (copied); every step
                           rebuilds V = V_xc + V_H + V_loc from its own rho, as the
                           density chares do (CP_Rho_RealSpacePlane.C:739-864) */
} oa_request;

typedef struct {
  long long tag;
  int status;           /* OA_OK converged, OA_ENOCONV step cap reached, else an error */
  int steps;
  double eelec, fnorm;  /* of the last step */
  double queue_ms;      /* submit -> admitted */
  double run_ms;        /* admitted -> done */
  double total_ms;      /* submit -> done */
  size_t hbm_bytes;     /* device memory the request held */
  int fp64_steps;       /* steps run in FP64 after a mixed-precision switch */
  int reused;           /* 1: the request ran on a pooled context of its shape (no
                           device allocation at admission) */
} oa_result;

/* HBM one context of this system needs (what admission charges) */
size_t oa_ctx_bytes(const oa_system *sys, const oa_options *opt);

/* hbm_budget_bytes: device memory the server may hold for admitted
 * requests; max_active: admitted requests at most (streams in flight). */
oa_status oa_server_create(int device, size_t hbm_budget_bytes, int max_active,
                           oa_server **out);
/* Queue a request (inputs copied).  OA_EINVAL if it can never fit the
 * budget or its layout is invalid. */
oa_status oa_server_submit(oa_server *s, const oa_request *req);
/* Move up to max finished results into res; returns how many. */
int oa_server_poll(oa_server *s, oa_result *res, int max);
/* Block until every submitted request has finished (results stay queued
 * for poll). */
oa_status oa_server_drain(oa_server *s);
/* Withdraw every request not finished yet: queued ones finish at once,
 * admitted ones after their step in flight, all with OA_ECANCELED (their
 * results carry the steps run so far).  Requests submitted later run
 * normally.  No counterpart in the reference (its jobs run to completion);
 * for a caller that bounds a run by time. */
oa_status oa_server_cancel(oa_server *s);
/* [0] admitted now, [1] queued, [2] finished, [3] HBM bytes held,
 * [4] peak admitted, [5] peak HBM bytes, [6] steps run, [7] submitted */
oa_status oa_server_stats(oa_server *s, double stats[8]);
oa_status oa_server_destroy(oa_server *s);

#ifdef __cplusplus
}
#endif

#endif /* OA_SERVE_H */
