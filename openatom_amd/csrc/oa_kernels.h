// oa_kernels.h -- host-side launch interface of the device kernels.
#pragma once

#include "oa_common.h"

namespace oa {

// Raise fn's dynamic-LDS limit on the current device to at least `bytes`
// (once per kernel, device and larger request; thread-safe).  oa_attr.hip
hipError_t grant_lds(const void* fn, size_t bytes);
// compute units of the current device (read once per device; thread-safe)
int cu_count();

// Everything the FFT chain needs for one system (device pointers).
struct FftCtx {
  Geom g;
  FftPlan px, py, pz;
  int lpb;    // z-lines per workgroup in the z passes
  int kc;     // kx columns per workgroup in the y passes
  int npair;  // pairs of real rows per workgroup in the x passes
  const double2* psi;     // [nst][ld] stored coefficients (FP64 master copy)
  double2* force;         // [nst][ld]
  const int* blk_pt0;     // [nblk+1] first point of each z-line block
  const int* pt_code;     // [npts] (kx0 << 31) | (line_in_block << 12) | z
  const int* line_dst;    // [nlines] kx * ny + (ky mod ny)
  const int* kylo;        // [nplane_x] populated ky range per kx column
  const int* kyhi;
  const double2 *twx, *twy, *twz;  // exp(-2 pi i j / n)
  const double* g2;       // [npts] |G|^2
  const double* occ;      // [nst]
  const double* vks;      // [N] V(r)
  void* zbuf;             // [nst][nz][nplane_x][ny] complex T
  void* xbuf;             // [nst][nz][ny][nplane_x] complex T
  void* psir;             // [nst][nz][ny][nx] T
  double* rho;            // [N]
  double* eke_part;       // [nst][nblk]
  double inv_vol, ecut_ry, scale;
  // fused plane kernels
  int kcp;                // kx columns per y-FFT batch
  int pcp;                // pairs of rows per x-FFT batch
  int cs;                 // states per workgroup
  int nchunk;             // state chunks (= rho partials)
  int plane_cfg;          // threads / register-slot variant
  double* rho_part;       // [nchunk][N]
  // two-factor engine
  const int4* lines;      // [nlines] (first point, length, first kz, kx*ny + ky mod ny)
  // fused two-factor plane passes
  void* zout;             // [nst][nz][nplane_x][ny] complex T: R->G plane output
  int nchunk_r;           // state chunks of the R->G plane pass ...
  int cs_r;               // ... and states per chunk
  int psir_state;         // state whose psi(r) the G->R plane pass writes to psir
  int prune;              // plane passes pruned to |kx|, |ky| <= (N - 1) / 4
  int prune_z;            // z passes pruned to |kz| <= (nz - 1) / 4
  // plane engine: psi(r) written by the G->R pass and read back by the R->G
  // pass, [nst][nz][plane3_kper] complex T in the x pass's lane order; null:
  // the R->G pass recomputes psi(r) from zbuf
  void* psk;
};

// two-factor register FFT engine (oa_fft2_*.hip): kind 0 z fwd, 1 y fwd,
// 2 x c2r + rho partials, 3 V psi + x r2c, 4 y bwd, 5 z bwd + kinetic/EKE;
// n is the length of the kind's axis
hipError_t fft2_launch(int kind, int n, bool fp32, const FftCtx& c, hipStream_t st);
int fft2_ntr(int n);  // transforms per workgroup for length n, 0 if not compiled

// fused y+x plane passes on the two-factor engine (oa_plane3.h), square
// planes nx == ny == n: kind 0 G->R + rho partials, 1 R->G (V psi, x and
// y back to zout), 2 G->R writing psi(r) of c.psir_state, 3 R->G that also
// leaves the rho partials of its c.nchunk_r chunks (V independent of rho)
hipError_t plane3_launch(int kind, int n, bool fp32, const FftCtx& c, hipStream_t st);
size_t plane3_smem(int n, int nplane_x, bool fp32);  // 0 if n is not compiled
// complex elements per state and z-plane of the kept psi(r) (FftCtx::psk)
size_t plane3_kper(int n, int nplane_x, bool fp32);  // 0 if n is not compiled
// row parts of the G->R grid and resident workgroups per CU of both passes
hipError_t plane3_info(int n, int nplane_x, bool fp32, bool prune, int* rowparts, int* occ_g2r,
                       int* occ_r2g, int* rho_ok);

// fused y+x plane passes (one z-plane, a chunk of states per workgroup)
hipError_t plane_fwd(const FftCtx& c, bool fp32, hipStream_t st);
hipError_t plane_bwd(const FftCtx& c, bool fp32, hipStream_t st);
// rho = sum of the chunk partials / Omega (accumulate: rho += that);
// with_v: also partial sums of V rho into dpart[ndpart]
hipError_t rho_reduce(const FftCtx& c, double* dpart, int ndpart, hipStream_t st,
                      bool with_v = true, bool accumulate = false);
size_t plane_smem(const FftCtx& c, bool fp32);
int plane_config_for(int plane_points);  // -1 if the plane is too large
int plane_threads(int cfg);

// stage: 0 z-gather+FFT, 1 y-FFT fwd, 2 x c2r + rho, 3 V psi + x r2c,
//        4 y-FFT bwd, 5 z-FFT + pack + kinetic force / EKE partials
enum { FFT_ZFWD = 0, FFT_YFWD, FFT_XC2R, FFT_XR2C, FFT_YBWD, FFT_ZBWD, FFT_NSTAGE };
hipError_t fft_stage(const FftCtx& c, bool fp32, int stage, hipStream_t st);
size_t fft_max_smem(const FftCtx& c, bool fp32);

// Device-resident per-step results (copied to the host once per step).
struct StepDev {
  double eke, eloc, fnorm2;
  double ns_err;
  int ns_iter;      // >0 converged at this iteration, <0 not converged
  int ns32_iter;    // FP32 iterations run before the FP64 ones (mixed Newton-Schulz)
  double resid_max, resid_rms;
  double ehart, exc, muxc;  // V_KS from rho: Hartree, exchange-correlation, sum V_xc rho dV
  double enl;               // non-local KB energy
};

// ---- non-local KB pseudopotential, l = 0 (oa_nonlocal.hip) ----
struct NlDev {
  int natoms, nspecies, nspline, npts, ldr;  // ldr: reals per table row (>= 2 npts)
  double gmin, dg, y00;                      // spline grid; y00 = 1/sqrt(4 pi)
  double hmati[9];
  const int *kx, *ky, *kz;  // [npts] stored k-vectors
  const double* pos;        // [natoms][3]
  const int* species;       // [natoms]
  const double* proj;       // [nspecies][4][nspline]
  const double* proj_g0;    // [nspecies]
};
// sf rows q * natoms + I: q = 0 the conjugated structure factor times the
// projector, q = 1..3 its derivative rows (d/dR_x, y, z); [4 natoms][ldr]
hipError_t nl_sf_build(const NlDev& d, double* sf, hipStream_t st);
// Mt[I][s] = 2 vn[I] Z[s][I]; *enl = sum_s occ[s] sum_I vn[I] Z[s][I]^2
hipError_t nl_energy(const double* Z, int ldz, int n, int natm, const double* vn,
                     const double* occ, double* Mt, double* enl, hipStream_t st);
// f[I][q] = -sum_s 2 vn[I] Z[s][I] dZ_q[s][I], Z: [n][4 natm]
hipError_t nl_atom_forces(const double* Z, int n, int natm, const double* vn, double* f,
                          hipStream_t st);

// ---- V_KS from rho (oa_vks.hip) ----
// Dense real 3D transform of rho restricted to the density sphere's box:
// kx in [0, nkx), ky in [-kyr, kyr] (nky = 2 kyr + 1), all kz.
struct VksArgs {
  int nx, ny, nz;
  int nkx, kyr, nky;
  int pr, cb, lb;          // pair-rows / kx columns / z-lines per workgroup
  FftPlan px, py, pz;
  const double2 *twx, *twy, *twz;  // exp(-2 pi i j / n)
  double hmati[9];
  double vol, ecut4;       // Omega; density cutoff 4 ecut_ry on |G|^2
  const double* rho;       // [nz][ny][nx]
  double* v;               // [nz][ny][nx] out: V_KS
  double2* hx;             // [nz][nkx][ny] after the x pass
  double2* hl;             // [nz][nkx][nky] after the y pass
  const double2* vloc;     // [nkx][nky][nz] V_loc(G), zero off the sphere
  double* part_x;          // [nbx][3] exchange-correlation partial sums
  double* part_l;          // [nbl][2] Hartree and E_ext partial sums
};
// workgroup counts of the x and line passes
int vks_nbx(const VksArgs& a);
int vks_nbl(const VksArgs& a);
size_t vks_smem(const VksArgs& a);  // largest dynamic LDS of the passes
hipError_t vks_build(const VksArgs& a, StepDev* res, hipStream_t st);
// V_loc(G) of fixed ions on the line box from per-species spline tables
struct PseudoDev {
  int natoms, nspecies, nspline;
  const double* pos;       // [natoms][3]
  const int* species;      // [natoms]
  double gmin, dg;
  const double* vps;       // [nspecies][4][nspline]
  const double* vps_g0;    // [nspecies]
};
hipError_t vloc_build(const VksArgs& a, const PseudoDev& p, double2* vloc, hipStream_t st);

// ---- dense linear algebra (oa_gemm.hip) ----
// part[sp][i][j] = sum_{k in slice sp} w(k) A[i][k] B[j][k]; A, B are [n][lda]
// rows of K reals (lda a multiple of 4, rows 32-byte aligned); w(k) = w_lo
// for k < k_lo else w_hi.  Every 128 x 128 output block is split into nslice
// K slices (slices never straddle k_lo).  sym: B == A, only the 16 x 16
// tiles with i/16 <= j/16 are computed.
// C[i][j] = sum_k w(k) A[i][k] B[j][k], i < n, j < nb (nb == n when sym)
struct GramArgs {
  const double* A;
  const double* B;
  int n, lda, K, k_lo;
  double w_lo, w_hi;
  int sym;
  int nslice;
  double* part;  // [nslice][n][nb]
  int nb;        // rows of B
};
// K slices per output block for about `target` workgroups in all
int gram_slices(int n, int K, int k_lo, int sym, int target);
hipError_t gram_f64(const GramArgs& a, hipStream_t st);
// C = alpha * sum_sp part[sp]; mirrored from the computed tiles if sym
hipError_t gram_reduce(const double* part, int nslice, int n, int nb, int sym, double alpha,
                       double* C, hipStream_t st);

// Y[j][k] = beta * Y[j][k] + s(k) * sum_{i < ni} M[i][j] X[i][k]; s(k) = s_lo
// for k < k_lo else s_hi.  X: [ni][ldx], Y: [n][ldx] (Y may not alias X),
// M: [ni][n]; ni = 0 means ni = n.
// With sd_psi set (n <= 128 only) the epilogue also applies the steepest-
// descent update to the rows of X: sd_psi[j][k] += dtm[k/2] * Ynew[j][k]
// (sd_psi == X), and writes per-workgroup partials of sum Ynew^2 to fpart.
// rotate_bf16 only: Xadd (may be null, not Y) is added to the result,
// Y = beta Y + s(k) M^T X + Xadd, and m_minus_i uses M - I for M (with
// Xadd = X: the update form X + X (M - I) of X M).
struct RotArgs {
  const double* X;
  const double* M;
  double* Y;
  int n, ldx, K, k_lo;
  double s_lo, s_hi, beta;
  double* sd_psi;
  const double* dtm;
  double* fpart;
  int ni;
  const double* Xadd;
  int m_minus_i;
};
hipError_t rotate_f64(const RotArgs& a, hipStream_t st);
int rotate_nwg(int n, int K);  // workgroups of a rotate launch (fpart entries)

// ---- split-BF16 products of the mixed-precision step (oa_gemm_bf16.hip) ----
// Each FP64 operand is split into three BF16 terms; the six term products
// with i + j <= 2 are summed on the matrix cores in FP32 (error ~ FP32
// accumulation over one K slice or over the inner dimension).
// gram_bf16: the product of gram_f64 with FP32 partials part[nslice][n][nb]
// (every block whole, also when sym), summed by gram_reduce_f32.
hipError_t gram_bf16(const GramArgs& a, float* part, hipStream_t st);
hipError_t gram_reduce_f32(const float* part, int nslice, int n, int nb, int sym, double alpha,
                           double* C, hipStream_t st);
// rotate_bf16: the product of rotate_f64 for n, ni <= 128 and ldx a multiple
// of 64 (no SD epilogue); fpart, if set, gets per-workgroup sums of Ynew^2
hipError_t rotate_bf16(const RotArgs& a, hipStream_t st);
int rotate_bf16_nwg(int K);  // workgroups of a rotate_bf16 launch (fpart entries)

// Newton-Schulz for T = (S/2)^{-1/2} with device-side convergence test
// (ortho.C:862-985).  Work: X[2], T[2] (n*n each), errpart[(max_iter+1) * nwg].
struct NsArgs {
  const double* S;
  double* X0; double* X1;
  double* T0; double* T1;
  double* C;
  double* Tout;
  double* errpart;
  StepDev* res;
  int n, max_iter;
  double tol;
  int coop;  // 0: two launches per iteration; 1: one cooperative launch over 32 x 32
             // tiles (two grid barriers per iteration); 2, 3, 4: one cooperative launch
             // over 16-row blocks, each split over 1, 2 or 4 workgroups by column
             // tiles (one barrier per iteration); kNsOneKind: one workgroup, a plain
             // launch, for n <= kNsOneMax (two launches per iteration above)
};
constexpr int kNsOneKind = 5;
constexpr int kNsOneMax = 64;
hipError_t ns_run(const NsArgs& a, hipStream_t st);
// FP32 work space of the mixed-precision Newton-Schulz: X and T (double
// buffered) and C, n x n floats each
struct Ns32Bufs {
  float* X0; float* X1;
  float* T0; float* T1;
  float* C;
};
// Newton-Schulz with its first iterations in FP32 on the matrix cores: FP32
// iterations until the rms change of T is <= kNs32Switch (at most
// a.max_iter - 1 of them), then FP64 iterations from that iterate to a.tol,
// a.max_iter iterations in all; two launches per iteration.
// res->ns32_iter holds the FP32 count, res->ns_iter the FP64 one (negative
// when the tolerance was not met).
constexpr double kNs32Switch = 1e-4;
hipError_t ns_run_mixed(const NsArgs& a, const Ns32Bufs& b, hipStream_t st);
bool ns_coop_possible(int n, int device);
bool ns_rows_possible(int n, int split, int device);
inline int ns_rows_split(int coop) { return 1 << (coop - 2); }  // coop 2, 3, 4
constexpr int kNsRowsMaxSplit = 4;

// Conjugate-gradient state of a minimisation, on the device
// (CP_State_GSpacePlane.C:120-136, 2646-2671)
struct CgDev {
  double fovlap, fovlap_old, fmag, fmag_old, gamma;
  int iter, num_reset;
};

// ---- pointwise G-space kernels (oa_pointwise.hip) ----
// gamma of the new direction from res->fnorm2 (one thread); the state advances
hipError_t cg_gamma(const StepDev* res, CgDev* cg, hipStream_t st);
// cg = gamma cg + F where g2 <= ecut; psi += dtm cg (cp_min_CG.C:55-105)
hipError_t cg_update(double2* psi, double2* cgdir, const double2* f, const double* dtm,
                     const double* g2, double ecut, int nst, int npts, int ld, const CgDev* cg,
                     hipStream_t st);
// psi += dtm[p] * F (cp_min_std.C:56-86); accumulates |F|^2 partials.
// twin[p] (may be null): -1, or -2 for a redundant kx=0 point (not
// written), or the redundant twin of p, which receives conj(new psi[p])
// (the redundancy sync, CP_State_GSpacePlane.C:2862-3061)
hipError_t sd_update(double2* psi, const double2* f, const double* dtm, const int* twin, int nst,
                     int npts, int ld, double* fpart, int nfpart, hipStream_t st);
// redundant kx=0 points <- conj(partner) (CP_State_GSpacePlane.C:2862-3061)
hipError_t red_sync(double2* psi, const int* red, const int* partner, int nred, int nst, int ld,
                    hipStream_t st);
// sum of parts -> res field (deterministic single-workgroup reduction)
hipError_t reduce_to(const double* part, int n, double scale, double* out, hipStream_t st);
// sum_r V rho partials
hipError_t dot_parts(const double* a, const double* b, size_t n, double* part, int nparts,
                     hipStream_t st);
// max|C/2 - I| and rms over an n*n matrix -> res->resid_max/resid_rms
hipError_t resid_stats(const double* S, int n, StepDev* res, hipStream_t st);
// zero-fill (graph safe replacement for memset)
hipError_t fill_zero(void* p, size_t bytes, hipStream_t st);

}  // namespace oa
