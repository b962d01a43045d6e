/*
 * oa_step.h -- C ABI of the MI355X (gfx950) CP electronic-step library
 * (openatom_amd/liboa_step.so).
 *
 * The library replaces, on one GPU, the hot path of OpenAtom's Gamma-point
 * electronic step.  Each entry point names the reference interface it stands
 * in for (paths relative to the OpenAtom tree).  All arithmetic at the
 * boundary is FP64; plain pointers and sizes only.
 *
 * Conventions (the reference's own):
 *   - states carry norm 2 and occupation 1 (gen_wave.C:699-710,
 *     parse.C:283-284), so S = Psi^H Psi is about 2I and Ortho works on S/2;
 *   - G-space coefficients are complex doubles, interleaved (re, im),
 *     [nstates][npts] in the order of the stored k-vectors: the kx=0 plane
 *     first (its redundant half ky<0 or ky=0,kz<0 included), points of one
 *     (kx,ky) line contiguous with kz ascending
 *     (stateSlab.C:285-380, util.C:1364-1397);
 *   - real-space arrays are [Nz][Ny][Nx] doubles, x fastest;
 *   - G->R is psi(r) = sum_G c(G) exp(-i G.r) unnormalised
 *     (fftCache.C:1133-1243); the force from V is
 *     F(G) = -(2/N) sum_r V psi exp(+i G.r), doubled off the kx=0 plane
 *     (CP_State_GSpacePlane.C:1607-1612, stateSlab.C:240-272).
 *
 * Ownership: host pointers are borrowed for the duration of a call
 * (copy-in / copy-out).  Device state belongs to the context.  A context is
 * bound to one HIP stream and is not thread-safe; many contexts may share a
 * process and a GPU.  Errors are status codes, never aborts; the message of
 * the last error is available from oa_last_error().
 */
#ifndef OA_STEP_H
#define OA_STEP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oa_ctx oa_ctx;

typedef enum {
  OA_OK = 0,
  OA_EINVAL = 1,   /* bad argument or layout (replaces the CkExit layout checks) */
  OA_EHIP = 2,     /* HIP runtime error */
  OA_ENOMEM = 3,   /* device allocation failed */
  OA_ENOCONV = 4,  /* Newton-Schulz missed its tolerance (ortho.C:129-133) */
  OA_ESTATE = 5,   /* call out of order (the reference's "Flow of Control" checks) */
  OA_ECANCELED = 6 /* serving: the request was withdrawn by oa_server_cancel */
} oa_status;

typedef enum { OA_FP64 = 0, OA_MIXED = 1 } oa_precision;

typedef struct {
  double hmat[9];      /* cell matrix, bohr, PINY flat order (hmat[1..9]) */
  int grid[3];         /* Nx, Ny, Nz (state FFT grid) */
  double ecut_ry;      /* wavefunction cutoff: |G|^2 <= ecut_ry */
  int nstates;
  const double *occ;   /* [nstates] occupations or NULL (all 1) */
  int npts;            /* stored points; 0 => enumerate from ecut/hmat */
  const int *kx, *ky, *kz; /* [npts] stored k-vectors or NULL */
} oa_system;

typedef struct {
  int precision;       /* oa_precision */
  int ns_max_iter;     /* invsqr_max_iter (ortho.h:86-87) */
  double ns_tol;       /* invsqr_tolerance */
  double dt;           /* SD time step, a.u. (gentimeinfo->dt) */
  double tau_mass;     /* cp_mass_tau_def, fs (default 25) */
  double ecut_mass;    /* cp_mass_cut_def (default 2) */
  int use_lambda;      /* project F with Lambda (cp_min_opt = 1) */
  int use_cg;          /* conjugate gradient instead of steepest descent (cp_min_cg = 1;
                          CP_integrate_min_CG, cp_min_CG.C:55-105, with the reset rule of
                          CP_State_GSpacePlane.C:2646-2671); needs use_lambda.  oa_psi_set
                          restarts the direction. */
  int want_resid;      /* compute ||Psi^H Psi/2 - I|| after each step */
  int fft_mode;        /* 0 (default): the fastest engine compiled for the grid (two-factor
                          fused planes, else two-factor per pass, else Stockham per pass);
                          1: Stockham LDS kernels, one per pass; 2: Stockham fused planes;
                          3: two-factor register engine, one kernel per pass;
                          4: two-factor fused y+x planes, psi(r) recomputed on chip
                          (square x-y planes only; OA_EINVAL where not compiled) */
  int use_graph;       /* 1: oa_step replays its kernel chain from HIP graphs (captured on
                          the second call; the Newton-Schulz launch stays outside); stage
                          timings then read: [0] everything before Newton-Schulz, [8]
                          Newton-Schulz, [9] Psi T (+ residual), [12] whole step */
  int ns_mode;         /* 0: one cooperative launch when all workgroups fit (default; the
                          autotuner picks among the forms below that fit),
                          1: two launches per iteration with a device-side early exit,
                          2: one cooperative launch over 32 x 32 tiles,
                          3, 4, 5: one cooperative launch over 16-row blocks, each split
                          over 1, 2 or 4 workgroups by 16-column tiles (4 and 5 need
                          n > 16 and n > 48); 2..5 give OA_EINVAL where they do not fit;
                          6: no cooperative launch: one plain launch of one workgroup for
                          n <= 64, else as 1 (for contexts that share the device) */
  int vks_from_rho;    /* 1: each step builds V(r) = V_KS[rho] from its own density
                          (PZ-LDA exchange-correlation, Hartree, local pseudopotential of
                          oa_local_pseudo_set) between rho and V psi, as the density chares
                          do; 0 (default): V(r) is the input of oa_vks_set */
  int autotune;        /* 1: oa_create times candidate launch shapes on the device and keeps
                          the fastest: state chunks of the two plane passes, K slices of
                          the S and Lambda products, Newton-Schulz as one cooperative launch
                          or two launches per iteration.  Chunking changes the order of the
                          rho sums (rounding only).  0 (default): fixed heuristics */
  int gemm_split;      /* mixed precision only: products run as split-BF16 matrix-core
                          products (each FP64 operand rounded to FP32 and split exactly into
                          three BF16 terms, the six term products with i + j <= 2 summed in
                          FP32; K-slice partials summed in FP64), bit 0 (default on): the
                          Lambda gram Re Psi^H F; bit 1: S = Psi^H Psi; bit 2: Psi T, in the
                          update form Psi + Psi (T - I); bit 3: the Lambda projection
                          F -= w Psi Lambda.  Bits 2 and 3 apply up to 128 states (FP64
                          above).  The ||Psi^H Psi/2 - I|| residual, Newton-Schulz and the
                          energies stay FP64.  Ignored in FP64 contexts. */
  int ns_fp32;         /* mixed precision only: Newton-Schulz (S/2)^{-1/2} runs its first
                          iterations in FP32 on the matrix cores, until the rms change of T
                          is <= 1e-4 (at most ns_max_iter - 1 of them), then FP64 iterations
                          from that iterate to ns_tol, ns_max_iter iterations in all (two
                          launches per iteration; oa_step_out.ns_iters counts both).  T meets
                          the same tolerance as in FP64.  0 (default): FP64 throughout */
  int fft_full;        /* 1: the state FFT passes (two-factor z passes, fused planes) run
                          every transform in full; 0 (default): along each axis where every
                          stored |k| is <= (N - 1) / 4 (always so on a grid that holds the
                          density sphere) they skip the radix-pass inputs that are zero
                          outside the sphere and the outputs that are not kept
                          (rounding-level differences only) */
  int psir_keep;       /* fused plane engine with separate G->R and R->G passes (V built
                          from rho, or rho not fused into the R->G pass): 1: the G->R pass
                          also writes every state's psi(r) to the device (FP32 or FP64,
                          about nstates x N reals) and the R->G pass reads it back, as the
                          reference's RealSpacePlane chares hold psi(r) between the density
                          and V psi (CP_State_RealSpacePlane.C:349-355, 530-597); 0
                          (default): the R->G pass recomputes psi(r) from the z-pass output
                          (no psi(r) buffer; measured faster on MI355X at 100^3: the extra
                          write and read of psi(r) cost more than the y-FFT and x c2r they
                          replace).  Bit-identical results either way. */
} oa_options;

typedef struct {
  double eke;          /* kinetic energy (cp_eke.C:128) */
  double eloc;         /* (Omega/N) sum_r V rho; with vks_from_rho the local
                          pseudopotential energy E_ext (cp_hart_ext.C:289, 323) */
  double eelec;        /* eke + eloc (+ ehart + exc with vks_from_rho) */
  double fnorm;        /* sqrt(sum |F|^2) after projection */
  int ns_iters;        /* Newton-Schulz iterations used */
  double ns_err;       /* final rms change of T (ortho.C:117) */
  double resid_max;    /* max |(Psi^H Psi)/2 - I| after the step (if asked) */
  double resid_rms;
  double ms;           /* device time of the step, ms */
  double ehart;        /* vks_from_rho: Hartree energy (cp_hart_ext.C:207, 385), else 0 */
  double exc;          /* vks_from_rho: PZ-LDA exchange-correlation energy
                          (cp_pz_lda.C:143-167), else 0 */
  double enl;          /* non-local KB energy when oa_nonlocal_set was called
                          (CP_enl_atm_forc_calc, cp_nl_energy_forc.C:196-237),
                          else 0; included in eelec */
} oa_step_out;

/* Local pseudopotential of fixed ions for a vks_from_rho context: per-species
 * form factors v_s(|G|) as cubic splines on the uniform grid g_i = gmin + i dg,
 * in the layout of PINY's vps0..vps3 tables, evaluated as CP_get_vpsnow does
 * (cp_hart_ext.C:433-475): row i = min(max(floor((g - gmin)/dg) + 1, 1),
 * nspline) (1-based), h = g - ((i-1) dg + gmin),
 * v = ((c3 h + c2) h + c1) h + c0; vps_g0 is the G = 0 value (gzvps). */
typedef struct {
  int natoms, nspecies, nspline;
  const double *pos;     /* [natoms][3] Cartesian positions, bohr */
  const int *species;    /* [natoms] in 0 .. nspecies-1 */
  double gmin, dg;
  const double *vps;     /* [nspecies][4][nspline]: rows c0, c1, c2, c3 */
  const double *vps_g0;  /* [nspecies] */
} oa_local_pseudo;

/* Non-local Kleinman-Bylander pseudopotential of fixed ions, l = 0 channel,
 * by the reference's G-space method (its default, cp_nonloc_ees_opt off;
 * that method supports l = 0 only, control_vps_params.C:1776-1792).  The
 * radial projectors P_t(|G|) use the spline layout of oa_local_pseudo
 * (get_rad_proj, cp_struct_fact.C:268-300; proj_g0 is the G = 0 value,
 * gzvps0), vnorm_t the KB normalisation (vnorm_0).  With
 * SF_I(G) = P_t(|G|) exp(-i G.R_I) / sqrt(4 pi) and w = 1 on kx = 0, 2
 * elsewhere, per state s: Z_sI = sum_stored w Re(psi_s conj SF_I),
 * E_NL = sum_s occ_s sum_I (vnorm_t/Omega) Z_sI^2,
 * F_s -= w sum_I (2 vnorm_t/Omega) Z_sI SF_I and, for the ions,
 * f_I -= sum_s (2 vnorm_t/Omega) Z_sI dZ_sI/dR_I; as in the reference, the
 * occupation weights the energy but not the two forces. */
typedef struct {
  int natoms, nspecies, nspline;
  const double *pos;     /* [natoms][3] Cartesian positions of the KB atoms, bohr */
  const int *species;    /* [natoms] in 0 .. nspecies-1 */
  double gmin, dg;
  const double *proj;    /* [nspecies][4][nspline]: rows c0, c1, c2, c3 */
  const double *proj_g0; /* [nspecies] */
  const double *vnorm;   /* [nspecies] */
} oa_nonlocal;

/* default options: FP64, ns_tol 1e-15, ns_max_iter 30, dt 5, tau 25 fs,
 * ecut_mass 2, Lambda on, residual on, gemm_split 1 (Lambda split-BF16 when
 * the context is mixed precision) */
void oa_default_options(oa_options *opt);

/* Stored k-vectors of a cell and cutoff in the reference order:
 * setkvec3d_sm (set_cp_ewald.C:523-680) + flip_data_set
 * (para_grp_parse.C:178-285).  Pass NULL arrays to get the count. */
int oa_kvecs(double ecut_ry, const double hmat[9], int *kx, int *ky, int *kz);

/* Build a context: layout, FFT plans, device buffers.
 * Replaces the FFTcache constructor (fftCache.C:37-449), the
 * GStateSlab/RealStateSlab allocation (stateSlab.C:33-106) and
 * create_line_decomp_descriptor (util.C:1650-1930). */
oa_status oa_create(int device, const oa_system *sys, const oa_options *opt, oa_ctx **out);
oa_status oa_destroy(oa_ctx *ctx);
const char *oa_last_error(const oa_ctx *ctx);

/* sizes of the context's layout: npts, nlines, nplane_x, nkx0 */
oa_status oa_layout_info(const oa_ctx *ctx, int info[4]);
/* device memory the context holds (the fftw_malloc'd slabs and plan
 * scratch of GStateSlab/RealStateSlab/FFTcache, stateSlab.C:88-106) */
size_t oa_ctx_device_bytes(const oa_ctx *ctx);
/* launch shapes in use: [0] FFT engine (0 two-factor per pass, 1 Stockham
 * per pass, 2 Stockham fused y+x planes, 3 two-factor fused y+x planes),
 * [1] z-lines, [2] y columns, [3] x pair-rows per workgroup (or per round),
 * [4] y columns per batch (Stockham fused) or states per R->G chunk
 * (two-factor fused), [5] x pair-rows per batch (Stockham fused) or R->G
 * chunks (two-factor fused), [6] states per workgroup in the G->R x pass,
 * [7] state chunks (rho partials) */
oa_status oa_get_config(const oa_ctx *ctx, int cfg[8]);
/* the stored k-vectors in use */
oa_status oa_get_kvecs(const oa_ctx *ctx, int *kx, int *ky, int *kz);

/* coefficients in/out, [nstates][npts] complex (GStateSlab::packedPlaneData) */
oa_status oa_psi_set(oa_ctx *ctx, const double *psi);
oa_status oa_psi_get(oa_ctx *ctx, double *psi);
/* V_KS(r) in [Nz][Ny][Nx] (the VksMsg payload, CP_State_RealSpacePlane.C:530-597) */
oa_status oa_vks_set(oa_ctx *ctx, const double *v);

/* G->R for all states + density.  Replaces doStpFFTGtoR_Gchare,
 * doStpFFTGtoR_Rchare (fftCache.C:1133-1243), the G->R transpose
 * (CP_State_GSpacePlane.C:1449-1518, CP_State_RealSpacePlane.C:186-271)
 * and the density reduction (CP_State_RealSpacePlane.C:349-355, 450-517;
 * CP_Rho_RealSpacePlane.C:398-447, 630-650). */
oa_status oa_fft_g2r(oa_ctx *ctx);
/* The same for the states [s0, s0 + ns) only, the granularity of the
 * reference's per-state-plane calls doStpFFTGtoR_Gchare / _Rchare
 * (fftCacheSlab.h:397-402) and of each state's contribution to the density
 * reduction (CP_State_RealSpacePlane.C:450-517): rho = (first ? 0 : rho) +
 * the range's share.  A caller that covers all states in ranges (first = 1
 * on the first) holds the density of oa_fft_g2r; with vks_from_rho it then
 * builds V with oa_vks_build.  Fused plane engine only (else OA_EINVAL). */
oa_status oa_fft_g2r_states(oa_ctx *ctx, int s0, int ns, int first);
/* psi_s(r) of one state (RealStateSlab plane data) */
oa_status oa_psir_get(oa_ctx *ctx, int state, double *psir);
/* rho(r) [Nz][Ny][Nx] */
oa_status oa_density_get(oa_ctx *ctx, double *rho);

/* V psi + R->G + EKE for all states.  Replaces unpackVks
 * (CP_State_RealSpacePlane.C:530-597), doStpFFTRtoG_Rchare/_Gchare
 * (fftCache.C:1099-1123, 1252-1319), doIFFT (CP_State_GSpacePlane.C:1582-1627),
 * GStateSlab::addForces (stateSlab.C:240-272) and CPNONLOCAL::CP_eke_calc
 * (cp_eke.C:30-140).  *eke may be NULL. */
oa_status oa_vpsi_r2g(oa_ctx *ctx, double *eke);
/* The same for the states [s0, s0 + ns) (doStpFFTRtoG_Rchare / _Gchare per
 * state plane, fftCacheSlab.h:397-402), after oa_fft_g2r or
 * oa_fft_g2r_states covered them; writes their forces, *eke is their share
 * of the kinetic energy.  Fused plane engine only (else OA_EINVAL). */
oa_status oa_vpsi_r2g_states(oa_ctx *ctx, int s0, int ns, double *eke);
oa_status oa_force_get(oa_ctx *ctx, double *f);
oa_status oa_force_set(oa_ctx *ctx, const double *f);

/* Lambda = Re Psi^H F and F -= w Psi Lambda (asymmetric PairCalculator
 * forward/backward, ckPairCalculator.C:682-802, 1189-1510;
 * CP_State_GSpacePlane.C:2045-2225).  lambda may be NULL. */
oa_status oa_lambda(oa_ctx *ctx, double *lambda);

/* psi += dt/m F and the kx=0 redundancy sync
 * (CPINTEGRATE::CP_integrate_min_STD, cp_min_std.C:56-86; CP_create_mass,
 * cp_control_integrate.C:136-203; CP_State_GSpacePlane.C:2862-3061) */
oa_status oa_sd_update(oa_ctx *ctx, double *fnorm);

/* S = Psi^H Psi (reference scaling, ~2I).  Replaces sendPsi's kx=0 scaling
 * (CP_State_GSpacePlane.C:3118-3125) + symmetric
 * PairCalculator::multiplyForward (ckPairCalculator.C:682-802) and the
 * Ortho tile assembly (ortho.C:153-268).  S may be NULL. */
oa_status oa_overlap(oa_ctx *ctx, double *S);

/* T = (S/2)^{-1/2} by Newton-Schulz on the S of the last oa_overlap.
 * Replaces Ortho::do_iteration/step_2/step_3/tolerance_check/collect_error
 * (ortho.C:109-146, 862-985) and CPORTHOG::CP_orthoTransform
 * (cp_lowdin.C:24-46, 180-265).  OA_ENOCONV if max_iter is reached. */
oa_status oa_ortho_ns(oa_ctx *ctx, double tol, int max_iter, int *iters, double *err, double *T);
/* T = (S/2)^{-1/2} of a caller's S, no context: the static
 * CPORTHOG::CP_orthoTransform(double *T, const double *S, int nstate)
 * (class_cporthog.h:24, cp_lowdin.C:24-46).  S and T are host arrays, n x n,
 * row-major, S in the reference's scaling (about 2I for orthonormal states).
 * Runs get_iter_Tmat's Newton-Schulz iteration (cp_lowdin.C:180-265: tol
 * 1e-15 on the rms change of T, at most 10 iterations) on the current HIP
 * device, for every n (the reference diagonalises instead when n <= 64,
 * get_diag_Tmat, whose T is the same matrix to the iteration's tolerance).
 * OA_ENOCONV where the reference EXITs (cp_lowdin.C:247-253); T then holds
 * the last iterate.  Allocates its device work space per call. */
oa_status oa_ortho_transform(double *T, const double *S, int n);

/* oa_ortho_transform without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The work space, S and T live in a fixed array of the library
 * (OA_ORTHO_NOALLOC_MAX_N = 512: 14.7 MB per device, held until the
 * process ends), set up on the first call on each device.  OA_EINVAL for
 * n > OA_ORTHO_NOALLOC_MAX_N.  Same kernels and result as
 * oa_ortho_transform.  All calls on one device share that device's array
 * and it keeps the last call's contents, so calls that overlapped would
 * race: the library serialises them with a per-device lock. */
#define OA_ORTHO_NOALLOC_MAX_N 512
oa_status oa_ortho_transform_noalloc(double *T, const double *S, int n);

/* load a T for oa_rotate (the PairCalculator's orthoT input) */
oa_status oa_tmat_set(oa_ctx *ctx, const double *T);

/* Psi <- Psi T.  Replaces PairCalculator::bwMultiplyHelper
 * (ckPairCalculator.C:1189-1425) and unpackNewPsi/doNewPsi
 * (CP_State_GSpacePlane.C:3185-3447). */
oa_status oa_rotate(oa_ctx *ctx);

/* One minimisation step in the order of GSpaceDriver (gspace.ci:106-190):
 * G->R, rho, V psi, R->G, EKE and kinetic force, [Lambda], SD update,
 * redundancy sync, S, Newton-Schulz, Psi T.  out may be NULL (no host
 * synchronisation); oa_step_result() fetches it later. */
oa_status oa_step(oa_ctx *ctx, oa_step_out *out);

/* V_KS from rho on the device (vks_from_rho contexts).
 * oa_local_pseudo_set computes V_loc(G) = (1/Omega) sum_I v_I(|G|)
 * exp(-i G.R_I) once for the fixed ions: the structure-factor loop of
 * CP_hart_eext_calc (cp_hart_ext.C:186-330) with the ion positions of
 * CP_Rho_GHartExt (CP_Rho_GHartExt.C:461-552).  Without it V_loc = 0.
 * oa_vks_build turns the context's rho (after oa_fft_g2r) into
 * V(r) = V_xc(r) + sum_{|G|^2 <= 4 ecut_ry} [4 pi rho(G)/(G^2 Omega) + V_loc(G)]
 * exp(+i G.r), rho(G) = (Omega/N) sum_r rho(r) exp(-i G.r): the work of
 * CP_Rho_RealSpacePlane (energyComputation, fftRhoRtoRhoG, acceptHartVks:
 * CP_Rho_RealSpacePlane.C:739-864, 1365-1390), CP_Rho_GSpacePlane and
 * CP_exc_calc (cp_pz_lda.C:23-170); its energies land in the next
 * oa_step_result (ehart, exc, eloc = E_ext).  oa_vks_get copies V(r) out. */
oa_status oa_local_pseudo_set(oa_ctx *ctx, const oa_local_pseudo *lp);

/* Set (or replace) the non-local atoms of a context: builds the projector
 * table once on the device (CP_calc_Struct_Fact, cp_struct_fact.C:38-260,
 * which the reference's StructFactCache holds per plane); every following
 * oa_step adds E_NL to eelec and the non-local force to F before Lambda,
 * where the reference's ParticlePlane chares join doneNLForces
 * (CP_State_ParticlePlane.C:500-790: computeZ, reduceZ, getForces).
 * oa_nonlocal_get evaluates, for the current coefficients, Z ([nstates]
 * [natoms], may be NULL), E_NL (may be NULL) and the ion forces
 * ([natoms][3], may be NULL); oa_nonlocal_force adds the non-local force to
 * the context's F (after oa_vpsi_r2g, before oa_lambda). */
oa_status oa_nonlocal_set(oa_ctx *ctx, const oa_nonlocal *nl);
oa_status oa_nonlocal_get(oa_ctx *ctx, double *z, double *enl, double *fatoms);
oa_status oa_nonlocal_force(oa_ctx *ctx);

/* What the autotuner measured and chose (one line per knob; empty when
 * opt.autotune was 0).  No counterpart in the reference, whose launch
 * shapes are fixed by its chare decomposition; the string lives as long as
 * the context. */
const char *oa_tuning_log(const oa_ctx *ctx);
oa_status oa_vks_build(oa_ctx *ctx);
oa_status oa_vks_get(oa_ctx *ctx, double *v);
/* ||Psi^H Psi / 2 - I|| of the current coefficients (max-abs and rms over
 * the n x n entries): the orthonormality that Ortho's tolerance_check
 * (ortho.C:109-146) controls through T; computed on demand, not per step
 * unless opt.want_resid is set.  Either pointer may be NULL. */
oa_status oa_residual(oa_ctx *ctx, double *resid_max, double *resid_rms);
oa_status oa_sync(oa_ctx *ctx);
oa_status oa_step_result(oa_ctx *ctx, oa_step_out *out);

/* Device time (ms) per stage of the timed oa_step calls once
 * oa_set_timing(ctx, 1) (average over the last <= 64 steps), from HIP
 * events on the context's stream: [0] z gather+FFT, [1] y-FFT (fused
 * planes: y+x G->R), [2] x c2r + rho (fused planes: rho reduction),
 * [3] V psi + x r2c (fused planes: V psi, x and y R->G), [4] y-FFT back
 * (fused planes: empty), [5] z-FFT + pack + kinetic force,
 * [6] EKE/E_loc reductions + Lambda (for n <= 128 with the SD update in the
 * same pass), [7] SD (or its |F| reduction) + redundancy sync, [8] S,
 * [9] Newton-Schulz, [10] Psi T, [11] residual, [12] whole step.
 * Fills min(n, OA_NTIMINGS) entries.  oa_set_timing(ctx, 2) records only
 * [3] and [12] and reports 0 for the other entries: every recorded event
 * delays the stream by a few microseconds, so level 1 slows the step it
 * measures. */
#define OA_NTIMINGS 13
oa_status oa_timings(oa_ctx *ctx, double *ms, int n);
oa_status oa_set_timing(oa_ctx *ctx, int on);

#ifdef __cplusplus
}
#endif
#endif
