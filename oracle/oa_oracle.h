/*
 * oa_oracle.h -- CPU restatement of OpenAtom's Gamma-point electronic-state
 * step, used ONLY as the parity checker (tests/, __graft_entry__.smoke())
 * and as the host-CPU baseline (bench.py cpu_baseline leg).  Nothing in the
 * product links or calls this library.
 *
 * Conventions (restated from the reference, paths relative to the
 * reference tree):
 *   - states carry norm 2 and occupation 1
 *     (src_piny_physics_v1.0/interface/coords_cp/gen_wave.C:699-710);
 *   - stored points: kx=0 plane first, redundant half (ky<0 or ky=0,kz<0)
 *     first, then G=0, then the PINY half sphere
 *     (src_charm_driver/utility/para_grp_parse.C:178-285);
 *   - G->R: psi(r) = sum_G c(G) exp(-i G.r), unnormalised, via z, y complex
 *     FFTs (FFTW_FORWARD), conjugation and an x c2r
 *     (src_charm_driver/fft_slab_ctrl/fftCache.C:1133-1243);
 *   - R->G: x r2c, conjugation, y and z FFTW_BACKWARD, pack, x(-2/N)
 *     (fftCache.C:1099-1123, 1252-1319;
 *      src_charm_driver/cp_state_ctrl/CP_State_GSpacePlane.C:1582-1627).
 * Real-space arrays are [Nz][Ny][Nx] with x fastest.
 */
#ifndef OA_ORACLE_H
#define OA_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* calc_cutoff kmax (set_cp_ewald.C:176-197); hmati row-major 3x3 */
void oo_kmax(double ecut_ry, const double hmati[9], int kmax[3]);

/* setkvec3d_sm (set_cp_ewald.C:523-680) followed by the PINY G=0 append;
 * returns the count (half sphere + G=0).  ka/kb/kc may be NULL to count. */
int oo_kvec_half(double ecut_ry, const double hmati[9], const int kmax[3],
                 int *ka, int *kb, int *kc);

/* flip_data_set (para_grp_parse.C:178-285): half sphere (with G=0 last)
 * -> stored order with the redundant kx=0 half first.  Returns npts. */
int oo_flip(int n, const int *ka, const int *kb, const int *kc,
            int *kx, int *ky, int *kz);

/* one-shot helper: enumerate the stored points of a cell/cutoff.
 * Pass NULL arrays to count only. */
int oo_stored_kvecs(double ecut_ry, const double hmat[9],
                    int *kx, int *ky, int *kz);

/* 3x3 inverse of a row-major matrix; returns the determinant */
double oo_inv3(const double h[9], double hi[9]);

/* textbook mixed-radix complex DFT, in place, unnormalised;
 * sign = -1 (FFTW_FORWARD) or +1 (FFTW_BACKWARD) */
void oo_fft1d(int n, int sign, double *data /* interleaved re,im [n] */);

/* naive O(n^2) DFT with the same conventions (pins oo_fft1d) */
void oo_dft1d(int n, int sign, const double *in, double *out);

typedef struct oo_sys oo_sys;

/* Build the oracle's view of one system.  kx/ky/kz: stored points in the
 * reference order (see header comment).  occ may be NULL (all ones). */
oo_sys *oo_create(int nx, int ny, int nz, const double hmat[9],
                  double ecut_ry, int nstates, const double *occ,
                  int npts, const int *kx, const int *ky, const int *kz);
void oo_destroy(oo_sys *s);
int oo_nlines(const oo_sys *s);
int oo_nplane_x(const oo_sys *s);
int oo_nkx0(const oo_sys *s);
int oo_nkx0_red(const oo_sys *s);
double oo_volume(const oo_sys *s);
void oo_g2(const oo_sys *s, double *g2 /* [npts] */);

/* per-state G->R (doStpFFTGtoR_Gchare + doStpFFTGtoR_Rchare) */
void oo_g2r(const oo_sys *s, const double *c /* [npts] cplx */,
            double *psir /* [Nz][Ny][Nx] */);
/* all states: psir[nst][N], threads over states */
void oo_g2r_all(const oo_sys *s, const double *psi, double *psir);

/* rho(r) = (1/Omega) sum_s occ_s psi_s(r)^2
 * (CP_State_RealSpacePlane.C:349-355, CP_Rho_RealSpacePlane.C:501,630-650) */
void oo_density(const oo_sys *s, const double *psir, double *rho);

/* per-state V psi product + R->G + (-2/N) scale
 * (CP_State_RealSpacePlane.C:580-596 then doStpFFTRtoG_*, doIFFT).
 * Output is the doIFFT force before GStateSlab::addForces. */
void oo_vpsi_r2g(const oo_sys *s, int istate, const double *psir_s,
                 const double *v, double *f /* [npts] cplx */);
void oo_vpsi_r2g_all(const oo_sys *s, const double *psir, const double *v,
                     double *f);

/* GStateSlab::addForces weighting (stateSlab.C:240-272, no non-local part)
 * + CP_eke_calc (cp_eke.C:30-140).  Returns EKE summed over states. */
double oo_eke_force(const oo_sys *s, const double *psi, double *f);

/* Lambda = Re sum_stored conj(psi_i) F_j  (asym PairCalculator forward,
 * alpha = 1, ckPairCalculator.C:682-802 with cpaimd.C:2527) and
 * F_j -= w_p sum_i psi_i Lambda_ij, w = 1/2 on kx=0, 1 elsewhere
 * (CP_State_GSpacePlane.C:2045-2225, cp_min_opt = 1, LDA) */
void oo_lambda(const oo_sys *s, const double *psi, const double *f,
               double *lambda /* [n][n] */);
void oo_lambda_apply(const oo_sys *s, const double *psi,
                     const double *lambda, double *f);

/* fictitious masses (cp_control_integrate.C:136-203) */
void oo_cmass(const oo_sys *s, double tau_fs, double ecut_mass,
              double *cmass /* [npts] */);
/* steepest descent psi += dt/m F (cp_min_std.C:56-86) */
void oo_sd_update(const oo_sys *s, const double *cmass, double dt,
                  const double *f, double *psi);
/* redundant kx=0 points <- conj(unique partner)
 * (CP_State_GSpacePlane.C:2862-3061) */
void oo_red_sync(const oo_sys *s, double *psi);

/* Conjugate-gradient minimisation state (CP_State_GSpacePlane.C:120-136,
 * 2646-2671): iteration counter, resets since the last one, |F|^2 and
 * |F| of this and the previous step. */
typedef struct {
  int iter, num_reset;
  double fovlap, fovlap_old, fmag, fmag_old;
} oo_cg_state;
/* Take the new sum |F|^2 (acceptCgOverlap) and return gamma of the
 * conjugate direction: 0 on the first iteration or on a reset
 * (|F| > 1.05 |F|_old after >= 5 steps without reset, or |F| > 2 |F|_old),
 * else fovlap / fovlap_old. */
double oo_cg_gamma(oo_cg_state *st, double fovlap);
/* cg = gamma cg + F where |G|^2 <= ecut; psi += dt/m cg
 * (CPINTEGRATE::CP_integrate_min_CG, cp_min_CG.C:55-105) */
void oo_cg_update(const oo_sys *s, const double *cmass, double dt, double gamma,
                  const double *f, double *cg, double *psi);

/* S = 2 sum_k A_ik A_jk over stored points with kx=0 scaled by 1/sqrt2
 * (sendPsi CP_State_GSpacePlane.C:3118-3125 + symm PC forward alpha=2,
 * ckPairCalculator.C:682-802) */
void oo_overlap(const oo_sys *s, const double *psi, double *S /* [n][n] */);

/* Newton-Schulz T = (S/2)^-1/2 (cp_lowdin.C:180-265; ortho.C:109-146,
 * 862-985).  Returns iterations; *err = final rms step; -1 if max_iter hit
 * with err > tol. */
int oo_ns(int n, const double *S, double tol, int max_iter, double *T,
          double *err);

/* psi'_j = sum_i psi_i T_ij (bwMultiplyHelper, ckPairCalculator.C:1189-1425) */
void oo_rotate(const oo_sys *s, const double *psi, const double *T,
               double *psi_out);

/* E_loc = (Omega/N) sum_r V rho */
double oo_eloc(const oo_sys *s, const double *v, const double *rho);

/* ---- V_KS(r) from rho(r): the density chares of the reference ----
 * Half-grid G arrays are [Nz][Ny][Nx/2+1] complex (kz, ky stored modulo
 * the grid, kx >= 0); the density sphere is |G|^2 <= 4 ecut_ry
 * (cp_hart_ext.C:175, ecut4 = 8 ecut_Ha).                              */

/* Local pseudopotential: per-species cubic splines of the form factor on a
 * uniform |G| grid (PINY vps0..vps3 tables) and the G=0 values (gzvps). */
typedef struct {
  int natoms, nspecies, nspline;
  const double *pos;     /* [natoms][3] Cartesian, bohr            */
  const int *species;    /* [natoms], 0 .. nspecies-1              */
  double gmin, dg;       /* spline grid: g_i = gmin + i dg         */
  const double *vps;     /* [nspecies][4][nspline]: c0, c1, c2, c3 */
  const double *vps_g0;  /* [nspecies]                             */
} oo_local_pseudo;

/* Spline value of one species at |G| = g (CP_get_vpsnow,
 * cp_hart_ext.C:433-475), without the 1/Omega. */
double oo_vps_eval(const oo_local_pseudo *lp, int species, double g);

/* V_loc(G) = (1/Omega) sum_I v_I(|G|) exp(-i G.R_I) on the density sphere,
 * v(0) = vps_g0 / Omega at G = 0, zero elsewhere: the conjugated structure
 * factor sum of CP_hart_eext_calc (cp_hart_ext.C:186-330). */
void oo_vloc_g(const oo_sys *s, const oo_local_pseudo *lp, double *vloc /* half grid */);

/* PZ-LDA (CP_exc_calc, cp_pz_lda.C:23-170): vxc[i] = V_xc(rho[i]) and the
 * three sums of section V-VI: sums[0] = sum xfact rho, sums[1] =
 * sum cfact rho, sums[2] = sum vxc rho. */
void oo_pz_lda(size_t n, const double *rho, double *vxc, double sums[3]);

typedef struct { double ehart, exc, eext, muxc; } oo_vks_energies;

/* V_KS(r) = V_xc(r) + sum_{|G|^2 <= 4 ecut} [4 pi rho_G/(G^2 Omega) + V_loc(G)]
 * exp(+i G.r), rho_G = (Omega/N) sum_r rho(r) exp(-i G.r)
 * (CP_Rho_RealSpacePlane.C:739-864, 1365-1390; cp_hart_ext.C:168-330):
 * E_H = 1/2 sum_G w 4 pi |rho_G|^2/(G^2 Omega), E_ext = sum_G w Re(rho_G
 * conj V_loc(G)) (w = 1 on kx = 0, 2 elsewhere), E_xc = (3/4 ex + ec)
 * Omega/N.  vloc may be NULL (no ions). */
void oo_vks(const oo_sys *s, const double *rho, const double *vloc, double *v,
            oo_vks_energies *e);

/* ---- Non-local Kleinman-Bylander pseudopotential, l = 0 channel ----
 * The reference's G-space ("N^3") method, its default (cp_nonloc_ees_opt
 * off, set_sim_dict.C:626-630), which supports only the l = 0 channel
 * (control_vps_params.C:1776-1792).  Per non-local atom I of species t:
 *   SF_I(G) = P_t(|G|) y00 exp(-i G.R_I),  y00 = 1/sqrt(4 pi)
 *             (CP_calc_Struct_Fact, cp_struct_fact.C:150-256: the table is
 *             built with exp(+i G.R) and conjugated for the double pack)
 *   Z_sI    = sum_stored w_G Re(psi_s(G) conj SF_I(G)), w = 1 on kx = 0,
 *             2 elsewhere (CP_enl_matrix_calc, cp_nlmat.C:86-118)
 *   E_NL    = sum_s occ_s sum_I (vnorm_t/Omega) Z_sI^2
 *   F_s(G) -= w_G sum_I (2 vnorm_t/Omega) Z_sI SF_I(G)
 *   f_I    -= sum_s (2 vnorm_t/Omega) Z_sI dZ_sI/dR_I
 * (CP_enl_atm_forc_calc, cp_nl_energy_forc.C:155-243; CP_enl_force_calc,
 * cp_nl_energy_forc.C:29-86).  As in the reference, the occupation
 * weights the energy only, not the two forces.                          */
typedef struct {
  int natoms, nspecies, nspline;
  const double *pos;      /* [natoms][3] Cartesian, bohr            */
  const int *species;     /* [natoms], 0 .. nspecies-1              */
  double gmin, dg;        /* spline grid: g_i = gmin + i dg         */
  const double *proj;     /* [nspecies][4][nspline]: c0, c1, c2, c3 */
  const double *proj_g0;  /* [nspecies]: P at G = 0 (gzvps0)        */
  const double *vnorm;    /* [nspecies]: KB normalisation vnorm_0   */
} oo_nonlocal;

/* Radial projector of one species at |G| = g (get_rad_proj,
 * cp_struct_fact.C:268-300). */
double oo_nl_proj(const oo_nonlocal *nl, int species, double g);

/* Z[nst][natoms] and, if dz is not NULL, dZ/dR as dz[nst][3][natoms]. */
void oo_nl_zmat(const oo_sys *s, const oo_nonlocal *nl, const double *psi, double *z,
                double *dz);

/* E_NL of psi; f (complex [nst][npts], may be NULL) and fatm ([natoms][3],
 * may be NULL) receive the non-local forces, added to what they hold. */
double oo_nl_apply(const oo_sys *s, const oo_nonlocal *nl, const double *psi, double *f,
                   double *fatm);

typedef struct {
  double dt;          /* SD step (a.u.)                        */
  double tau_fs;      /* cp_mass_tau_def (fs), default 25      */
  double ecut_mass;   /* cp_mass_cut_def, default 2            */
  double ns_tol;      /* invsqr_tolerance                      */
  int ns_max_iter;    /* invsqr_max_iter                       */
  int use_lambda;     /* project forces with Lambda            */
  int use_cg;         /* conjugate gradient instead of SD      */
  oo_cg_state *cg_state;  /* CG only: persistent state         */
  double *cg;         /* CG only: [nst][npts] complex direction */
  int vks_from_rho;   /* V = V_KS(rho) of this step (oo_vks)   */
  const double *vloc; /* V_KS only: V_loc(G) half grid or NULL */
  double *vks_out;    /* V_KS only: the step's V(r), or NULL   */
  const oo_nonlocal *nl;  /* non-local KB part, or NULL        */
} oo_step_opts;

typedef struct {
  double eke, eloc, eelec;
  double fnorm;       /* sqrt(sum |F|^2) after projection        */
  int ns_iters;
  double ns_err;
  double ehart, exc;  /* V_KS only (eloc is then E_ext)          */
  double enl;         /* non-local energy (opts.nl), else 0      */
} oo_step_out;

/* One minimisation step on the host (gspace.ci:106-190 order; with
 * vks_from_rho, V_KS is built from this step's rho before V psi):
 * G->R, rho, V psi, R->G, EKE + kinetic force, [non-local force],
 * [Lambda], SD, red sync, S, NS, Psi.T.  psi is updated in place; rho/S/T/f may be NULL. */
int oo_step(const oo_sys *s, const oo_step_opts *o, const double *v,
            double *psi, double *rho, double *S, double *T, double *f,
            oo_step_out *out);

/* thread control for the CPU baseline */
void oo_set_threads(int n);
int oo_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
