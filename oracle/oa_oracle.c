/*
 * oa_oracle.c -- CPU restatement of the OpenAtom Gamma-point electronic
 * step.  TEST INFRASTRUCTURE: the parity checker and the host-CPU
 * baseline.  It is never linked into, loaded by or called from the
 * product library.  Every function cites the reference file:line it
 * restates (paths relative to the reference tree).
 *
 * Parity pins for each function live in tests/test_oracle.py (numpy.fft,
 * brute-force DFT, numpy eigh closed form, sum rules, invariants).
 */
#include "oa_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

#define TIME_CONV 0.0241888 /* piny_constants.h:58 */
#define CP_EMAGIC 0.00050   /* piny_constants.h:59 */

static int g_threads = 0;
void oo_set_threads(int n) { g_threads = n; }
int oo_get_threads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

static int imod(int a, int n) { int r = a % n; return r < 0 ? r + n : r; }

double oo_inv3(const double h[9], double hi[9]) {
  double det = h[0] * (h[4] * h[8] - h[5] * h[7]) -
               h[1] * (h[3] * h[8] - h[5] * h[6]) +
               h[2] * (h[3] * h[7] - h[4] * h[6]);
  double id = 1.0 / det;
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

/* |G|^2/2 for integer triple, restating the xk/yk/zk expressions of
 * set_cp_ewald.C:563-566 with 0-based hmati (PINY hmati[1..9]). */
static double half_g2(const double *hi, double a, double b, double c) {
  const double tpi = 2.0 * M_PI;
  double xk = (a * hi[0] + b * hi[1] + c * hi[2]) * tpi;
  double yk = (a * hi[3] + b * hi[4] + c * hi[5]) * tpi;
  double zk = (a * hi[6] + b * hi[7] + c * hi[8]) * tpi;
  return (xk * xk + yk * yk + zk * zk) * 0.5;
}

/* set_cp_ewald.C:176-197 (cp_on branch).  The internal cutoff is in
 * Hartree: ecut_ha = ecut_ry/2 (set_cp_ewald.C:168-171). */
void oo_kmax(double ecut_ry, const double hi[9], int kmax[3]) {
  double ecut = 0.5 * ecut_ry;
  for (int d = 0; d < 3; d++) {
    double t = sqrt(hi[d] * hi[d] + hi[d + 3] * hi[d + 3] + hi[d + 6] * hi[d + 6]);
    double temp = sqrt(ecut * 0.5) / (M_PI * t);
    int it = (int)temp;
    if (temp - (double)it > 0.99) it++;
    kmax[d] = it;
  }
}

/* set_cp_ewald.C:523-680: enumerate the half sphere tryme<=ecut in the
 * (ka, kb, kc) loop order, then the G=0 append (set_cp_ewald.C:673-676). */
int oo_kvec_half(double ecut_ry, const double hi[9], const int kmax[3],
                 int *ka_o, int *kb_o, int *kc_o) {
  double ecut = 0.5 * ecut_ry;
  int icount = 0, i;
  for (i = 1; i <= kmax[0]; ++i)
    if (half_g2(hi, i, 0, 0) > ecut) break;
  int kamax = i - 1;
  for (int ka = 0; ka <= kamax; ++ka) {
    int kbmin = (ka == 0) ? 0 : -kmax[1];
    for (i = kbmin; i <= 0; ++i)
      if (half_g2(hi, ka, i, 0) <= ecut) break;
    kbmin = i;
    for (i = 1; i <= kmax[1]; ++i)
      if (half_g2(hi, ka, i, 0) > ecut) break;
    int kbmax = i - 1;
    for (int kb = kbmin; kb <= kbmax; ++kb) {
      int kcmin = (ka == 0 && kb == 0) ? 1 : -kmax[2];
      for (i = kcmin; i <= 0; ++i)
        if (half_g2(hi, ka, kb, i) <= ecut) break;
      kcmin = i;
      for (i = 1; i <= kmax[2]; ++i)
        if (half_g2(hi, ka, kb, i) > ecut) break;
      int kcmax = i - 1;
      for (int kc = kcmin; kc <= kcmax; ++kc) {
        if (ka_o) { ka_o[icount] = ka; kb_o[icount] = kb; kc_o[icount] = kc; }
        icount++;
      }
    }
  }
  if (ka_o) { ka_o[icount] = 0; kb_o[icount] = 0; kc_o[icount] = 0; }
  return icount + 1;
}

/* para_grp_parse.C:178-285: prepend the mirror image of the kx=0 half
 * plane (reverse order, ky,kz negated), put G=0 after it. */
int oo_flip(int n, const int *ka, const int *kb, const int *kc,
            int *kx, int *ky, int *kz) {
  int nplane0 = 0;
  for (int i = 0; i < n; i++) if (ka[i] == 0) nplane0++;
  int npts = n + nplane0 - 1;
  if (!kx) return npts;
  /* mirrored redundant points: ind = nplane0 - i - 2 */
  for (int i = 0; i < nplane0 - 1; i++) {
    int ind = nplane0 - i - 2;
    kx[i] = ka[ind]; ky[i] = -kb[ind]; kz[i] = -kc[ind];
  }
  kx[nplane0 - 1] = ka[n - 1]; ky[nplane0 - 1] = kb[n - 1]; kz[nplane0 - 1] = kc[n - 1];
  for (int i = 0; i < n - 1; i++) {
    kx[nplane0 + i] = ka[i]; ky[nplane0 + i] = kb[i]; kz[nplane0 + i] = kc[i];
  }
  return npts;
}

int oo_stored_kvecs(double ecut_ry, const double hmat[9], int *kx, int *ky, int *kz) {
  double hi[9]; int kmax[3];
  oo_inv3(hmat, hi);
  oo_kmax(ecut_ry, hi, kmax);
  int n = oo_kvec_half(ecut_ry, hi, kmax, NULL, NULL, NULL);
  int *ka = malloc(sizeof(int) * n), *kb = malloc(sizeof(int) * n), *kc = malloc(sizeof(int) * n);
  oo_kvec_half(ecut_ry, hi, kmax, ka, kb, kc);
  int npts = oo_flip(n, ka, kb, kc, kx, ky, kz);
  free(ka); free(kb); free(kc);
  return npts;
}

/* ---------------- textbook mixed-radix DFT ---------------- */

/* exp(sign*2*pi*i*j/n), j < n */
static cplx *twiddles(int n, int sign) {
  cplx *t = malloc(sizeof(cplx) * n);
  for (int j = 0; j < n; j++) {
    double a = 2.0 * M_PI * (double)j / (double)n;
    t[j] = cos(a) + (double)sign * sin(a) * I;
  }
  return t;
}

static int smallest_factor(int n) {
  for (int p = 2; p * p <= n; p++) if (n % p == 0) return p;
  return n;
}

/* decimation in time: X[k] = sum_q W_n^{qk} Y_q[k mod m], n = p*m,
 * Y_q = DFT_m(x[q + p*j]).  tw is the table for the top size n0. */
static void fft_rec(int n, const cplx *in, int is, cplx *out, cplx *scr,
                    const cplx *tw, int n0) {
  if (n == 1) { out[0] = in[0]; return; }
  int p = smallest_factor(n), m = n / p;
  for (int q = 0; q < p; q++)
    fft_rec(m, in + (size_t)q * is, is * p, scr + (size_t)q * m, out, tw, n0);
  int step = n0 / n;
  for (int k = 0; k < n; k++) {
    cplx acc = 0;
    int km = k % m;
    for (int q = 0; q < p; q++) {
      int e = (int)(((long)q * k) % n);
      acc += tw[(size_t)e * step] * scr[(size_t)q * m + km];
    }
    out[k] = acc;
  }
}

typedef struct { int n; cplx *twf, *twb; } plan1d;

static void fft_plan(plan1d *pl, int n) { pl->n = n; pl->twf = twiddles(n, -1); pl->twb = twiddles(n, +1); }
static void fft_free(plan1d *pl) { free(pl->twf); free(pl->twb); }

/* in place; work must hold 3n */
static void fft_exec(const plan1d *pl, int sign, cplx *x, cplx *work) {
  int n = pl->n;
  cplx *out = work, *scr = work + n, *in = work + 2 * n;
  memcpy(in, x, sizeof(cplx) * n);
  fft_rec(n, in, 1, out, scr, sign < 0 ? pl->twf : pl->twb, n);
  memcpy(x, out, sizeof(cplx) * n);
}

void oo_fft1d(int n, int sign, double *data) {
  plan1d pl; fft_plan(&pl, n);
  cplx *w = malloc(sizeof(cplx) * 3 * n);
  fft_exec(&pl, sign, (cplx *)data, w);
  free(w); fft_free(&pl);
}

void oo_dft1d(int n, int sign, const double *in_, double *out_) {
  const cplx *in = (const cplx *)in_;
  cplx *out = (cplx *)out_;
  for (int k = 0; k < n; k++) {
    cplx acc = 0;
    for (int j = 0; j < n; j++) {
      double a = 2.0 * M_PI * (double)(((long)j * k) % n) / (double)n;
      acc += in[j] * (cos(a) + (double)sign * sin(a) * I);
    }
    out[k] = acc;
  }
}

/* ---------------- system ---------------- */

struct oo_sys {
  int nx, ny, nz, nst, npts, nlines, nplane_x, nkx0, nkx0_red;
  double hmat[9], hmati[9], vol, ecut_ry;
  double *occ;
  int *kx, *ky, *kz;
  int *loff, *llen, *lkx, *lky, *lkz0;
  int *partner; /* redundant point -> unique partner, -1 otherwise */
  double *g2;
  plan1d px, py, pz;
};

int oo_nlines(const oo_sys *s) { return s->nlines; }
int oo_nplane_x(const oo_sys *s) { return s->nplane_x; }
int oo_nkx0(const oo_sys *s) { return s->nkx0; }
int oo_nkx0_red(const oo_sys *s) { return s->nkx0_red; }
double oo_volume(const oo_sys *s) { return s->vol; }
void oo_g2(const oo_sys *s, double *g2) { memcpy(g2, s->g2, sizeof(double) * s->npts); }

static double full_g2(const double *hi, int a, int b, int c) { return 2.0 * half_g2(hi, a, b, c); }

oo_sys *oo_create(int nx, int ny, int nz, const double hmat[9], double ecut_ry,
                  int nst, const double *occ, int npts, const int *kx,
                  const int *ky, const int *kz) {
  oo_sys *s = calloc(1, sizeof(oo_sys));
  s->nx = nx; s->ny = ny; s->nz = nz; s->nst = nst; s->npts = npts;
  s->ecut_ry = ecut_ry;
  memcpy(s->hmat, hmat, sizeof(double) * 9);
  s->vol = fabs(oo_inv3(hmat, s->hmati));
  s->occ = malloc(sizeof(double) * nst);
  for (int i = 0; i < nst; i++) s->occ[i] = occ ? occ[i] : 1.0;
  s->kx = malloc(sizeof(int) * npts); s->ky = malloc(sizeof(int) * npts); s->kz = malloc(sizeof(int) * npts);
  memcpy(s->kx, kx, sizeof(int) * npts); memcpy(s->ky, ky, sizeof(int) * npts); memcpy(s->kz, kz, sizeof(int) * npts);
  s->g2 = malloc(sizeof(double) * npts);
  for (int i = 0; i < npts; i++) s->g2[i] = full_g2(s->hmati, kx[i], ky[i], kz[i]);
  /* lines: maximal runs of constant (kx,ky), kz ascending
   * (util.C:1364-1397 layout checks) */
  int nl = 0, kxm = 0;
  for (int i = 0; i < npts; i++) {
    if (i == 0 || kx[i] != kx[i - 1] || ky[i] != ky[i - 1]) nl++;
    if (kx[i] > kxm) kxm = kx[i];
  }
  s->nlines = nl; s->nplane_x = kxm + 1;
  s->loff = malloc(sizeof(int) * nl); s->llen = malloc(sizeof(int) * nl);
  s->lkx = malloc(sizeof(int) * nl); s->lky = malloc(sizeof(int) * nl); s->lkz0 = malloc(sizeof(int) * nl);
  int l = -1;
  for (int i = 0; i < npts; i++) {
    if (i == 0 || kx[i] != kx[i - 1] || ky[i] != ky[i - 1]) {
      l++; s->loff[l] = i; s->llen[l] = 0; s->lkx[l] = kx[i]; s->lky[l] = ky[i]; s->lkz0[l] = kz[i];
    }
    s->llen[l]++;
  }
  /* stateSlab.C:285-380 counts */
  s->nkx0 = 0; s->nkx0_red = 0;
  for (int i = 0; i < npts; i++) {
    if (kx[i] == 0) s->nkx0++;
    if (kx[i] == 0 && (ky[i] < 0 || (ky[i] == 0 && kz[i] < 0))) s->nkx0_red++;
  }
  /* partner map for the redundant half (sendRedPsi/unpackRedPsi) */
  s->partner = malloc(sizeof(int) * npts);
  for (int i = 0; i < npts; i++) s->partner[i] = -1;
  for (int i = 0; i < npts; i++) {
    if (!(kx[i] == 0 && (ky[i] < 0 || (ky[i] == 0 && kz[i] < 0)))) continue;
    for (int j = 0; j < npts; j++)
      if (kx[j] == 0 && ky[j] == -ky[i] && kz[j] == -kz[i]) { s->partner[i] = j; break; }
  }
  fft_plan(&s->px, nx); fft_plan(&s->py, ny); fft_plan(&s->pz, nz);
  return s;
}

void oo_destroy(oo_sys *s) {
  if (!s) return;
  free(s->occ); free(s->kx); free(s->ky); free(s->kz); free(s->g2);
  free(s->loff); free(s->llen); free(s->lkx); free(s->lky); free(s->lkz0); free(s->partner);
  fft_free(&s->px); fft_free(&s->py); fft_free(&s->pz);
  free(s);
}

/* ---------------- G -> R ---------------- */

void oo_g2r(const oo_sys *s, const double *c_, double *psir) {
  const cplx *c = (const cplx *)c_;
  int Nx = s->nx, Ny = s->ny, Nz = s->nz, nl = s->nlines, nh = Nx / 2 + 1;
  int nmax = Nx > Ny ? Nx : Ny; if (Nz > nmax) nmax = Nz;
  cplx *L = calloc((size_t)nl * Nz, sizeof(cplx));
  cplx *P = malloc(sizeof(cplx) * (size_t)Ny * nh);
  cplx *tmp = malloc(sizeof(cplx) * nmax), *w = malloc(sizeof(cplx) * 3 * nmax);
  /* expandGSpace (fftCache.C:460-507): line buffer index = kz mod Nz */
  for (int l = 0; l < nl; l++)
    for (int i = 0; i < s->llen[l]; i++)
      L[(size_t)l * Nz + imod(s->lkz0[l] + i, Nz)] = c[s->loff[l] + i];
  /* z FFT, FFTW_FORWARD (fftCache.C:1149-1158) */
  for (int l = 0; l < nl; l++) fft_exec(&s->pz, -1, L + (size_t)l * Nz, w);
  for (int z = 0; z < Nz; z++) {
    memset(P, 0, sizeof(cplx) * (size_t)Ny * nh);
    /* unpackFFT transpose (CP_State_RealSpacePlane.C:262-268) */
    for (int l = 0; l < nl; l++) P[(size_t)imod(s->lky[l], Ny) * nh + s->lkx[l]] = L[(size_t)l * Nz + z];
    /* y FFT on nplane_x columns, FFTW_FORWARD (fftCache.C:1180-1189) */
    for (int kx = 0; kx < s->nplane_x; kx++) {
      for (int y = 0; y < Ny; y++) tmp[y] = P[(size_t)y * nh + kx];
      fft_exec(&s->py, -1, tmp, w);
      for (int y = 0; y < Ny; y++) P[(size_t)y * nh + kx] = tmp[y];
    }
    /* conjugate (fftCache.C:1207-1213) then FFTW c2r:
     * out[x] = sum_k H[k] e^{+2 pi i k x/N}, H hermitian extension, imaginary
     * parts of H[0] (and H[N/2] for even N) ignored */
    for (int y = 0; y < Ny; y++) {
      cplx *row = P + (size_t)y * nh;
      for (int k = 0; k < Nx; k++) tmp[k] = 0;
      for (int k = 0; k < nh; k++) tmp[k] = conj(row[k]);
      tmp[0] = creal(tmp[0]);
      if (Nx % 2 == 0) tmp[Nx / 2] = creal(tmp[Nx / 2]);
      for (int k = 1; k < (Nx + 1) / 2; k++) tmp[Nx - k] = conj(tmp[k]);
      fft_exec(&s->px, +1, tmp, w);
      double *out = psir + ((size_t)z * Ny + y) * Nx;
      for (int x = 0; x < Nx; x++) out[x] = creal(tmp[x]);
    }
  }
  free(L); free(P); free(tmp); free(w);
}

void oo_g2r_all(const oo_sys *s, const double *psi, double *psir) {
  size_t N = (size_t)s->nx * s->ny * s->nz;
#pragma omp parallel for schedule(dynamic, 1) num_threads(oo_get_threads())
  for (int st = 0; st < s->nst; st++)
    oo_g2r(s, psi + (size_t)st * 2 * s->npts, psir + (size_t)st * N);
}

void oo_density(const oo_sys *s, const double *psir, double *rho) {
  size_t N = (size_t)s->nx * s->ny * s->nz;
  double inv = 1.0 / s->vol; /* probScale, CP_Rho_RealSpacePlane.C:138 */
#pragma omp parallel for schedule(static) num_threads(oo_get_threads())
  for (long r = 0; r < (long)N; r++) {
    double acc = 0.0;
    for (int st = 0; st < s->nst; st++) {
      double p = psir[(size_t)st * N + r];
      acc += p * p * s->occ[st]; /* wght_rho = occ * wght_kpt, wght_kpt = 1 */
    }
    rho[r] = acc * inv;
  }
}

/* ---------------- R -> G ---------------- */

void oo_vpsi_r2g(const oo_sys *s, int istate, const double *psir, const double *v, double *f_) {
  cplx *f = (cplx *)f_;
  int Nx = s->nx, Ny = s->ny, Nz = s->nz, nl = s->nlines, nh = Nx / 2 + 1;
  int nmax = Nx > Ny ? Nx : Ny; if (Nz > nmax) nmax = Nz;
  double occ = s->occ[istate];
  cplx *L = calloc((size_t)nl * Nz, sizeof(cplx));
  cplx *P = malloc(sizeof(cplx) * (size_t)Ny * nh);
  cplx *tmp = malloc(sizeof(cplx) * nmax), *w = malloc(sizeof(cplx) * 3 * nmax);
  for (int z = 0; z < Nz; z++) {
    for (int y = 0; y < Ny; y++) {
      const double *pr = psir + ((size_t)z * Ny + y) * Nx;
      const double *vr = v + ((size_t)z * Ny + y) * Nx;
      /* unpackVks product (CP_State_RealSpacePlane.C:580-588) */
      for (int x = 0; x < Nx; x++) tmp[x] = pr[x] * (vr[x] * occ);
      /* FFTW r2c: X[k] = sum_x f[x] e^{-2 pi i k x/N}, k <= N/2 */
      fft_exec(&s->px, -1, tmp, w);
      /* conjugate (fftCache.C:1272-1274) */
      for (int k = 0; k < nh; k++) P[(size_t)y * nh + k] = conj(tmp[k]);
    }
    /* y FFT FFTW_BACKWARD on nplane_x columns (fftCache.C:1294-1303) */
    for (int kx = 0; kx < s->nplane_x; kx++) {
      for (int y = 0; y < Ny; y++) tmp[y] = P[(size_t)y * nh + kx];
      fft_exec(&s->py, +1, tmp, w);
      for (int y = 0; y < Ny; y++) P[(size_t)y * nh + kx] = tmp[y];
    }
    /* transpose back to lines (sendFPsiToGSP / unpackIFFT) */
    for (int l = 0; l < nl; l++) L[(size_t)l * Nz + z] = P[(size_t)imod(s->lky[l], Ny) * nh + s->lkx[l]];
  }
  /* z FFT FFTW_BACKWARD, pack, x(-2/N) (fftCache.C:1099-1123,
   * CP_State_GSpacePlane.C:1607-1612) */
  double scale = -2.0 / ((double)Nx * Ny * Nz);
  for (int l = 0; l < nl; l++) {
    cplx *line = L + (size_t)l * Nz;
    fft_exec(&s->pz, +1, line, w);
    for (int i = 0; i < s->llen[l]; i++)
      f[s->loff[l] + i] = line[imod(s->lkz0[l] + i, Nz)] * scale;
  }
  free(L); free(P); free(tmp); free(w);
}

void oo_vpsi_r2g_all(const oo_sys *s, const double *psir, const double *v, double *f) {
  size_t N = (size_t)s->nx * s->ny * s->nz;
#pragma omp parallel for schedule(dynamic, 1) num_threads(oo_get_threads())
  for (int st = 0; st < s->nst; st++)
    oo_vpsi_r2g(s, st, psir + (size_t)st * N, v, f + (size_t)st * 2 * s->npts);
}

/* ---------------- G-space pointwise ---------------- */

double oo_eke_force(const oo_sys *s, const double *psi_, double *f_) {
  double eke_tot = 0.0;
  for (int st = 0; st < s->nst; st++) {
    const cplx *psi = (const cplx *)psi_ + (size_t)st * s->npts;
    cplx *f = (cplx *)f_ + (size_t)st * s->npts;
    double eke = 0.0;
    for (int i = 0; i < s->npts; i++) {
      /* addForces: kx>0 doubled (stateSlab.C:262-268), non-local absent */
      double w = (s->kx[i] == 0) ? 1.0 : 2.0;
      if (s->kx[i] != 0) f[i] *= 2.0;
      /* CP_eke_calc (cp_eke.C:72-120): wght_tot = occ*wght_kpt*w */
      double wt = s->occ[st] * w;
      if (s->g2[i] <= s->ecut_ry) {
        f[i] -= psi[i] * (wt * s->g2[i]);
        double m2 = creal(psi[i]) * creal(psi[i]) + cimag(psi[i]) * cimag(psi[i]);
        eke += (wt * s->g2[i]) * m2;
      } else {
        f[i] = 0.0;
      }
    }
    eke_tot += eke / 2.0; /* cp_eke.C:128 */
  }
  return eke_tot;
}

void oo_lambda(const oo_sys *s, const double *psi, const double *f, double *lam) {
  int n = s->nst; size_t K = 2 * (size_t)s->npts;
#pragma omp parallel for schedule(dynamic, 1) num_threads(oo_get_threads())
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      const double *a = psi + (size_t)i * K, *b = f + (size_t)j * K;
      double acc = 0.0;
      for (size_t k = 0; k < K; k++) acc += a[k] * b[k];
      lam[(size_t)i * n + j] = acc;
    }
}

void oo_lambda_apply(const oo_sys *s, const double *psi, const double *lam, double *f) {
  int n = s->nst; size_t K = 2 * (size_t)s->npts;
#pragma omp parallel for schedule(static) num_threads(oo_get_threads())
  for (long k = 0; k < (long)K; k++) {
    int p = (int)(k / 2);
    double w = (s->kx[p] == 0) ? 0.5 : 1.0; /* ws = 1/2, wd = 2/2 */
    for (int j = 0; j < n; j++) {
      double acc = 0.0;
      for (int i = 0; i < n; i++) acc += psi[(size_t)i * K + k] * lam[(size_t)i * n + j];
      f[(size_t)j * K + k] -= w * acc;
    }
  }
}

void oo_cmass(const oo_sys *s, double tau_fs, double ecut_mass, double *cmass) {
  double tau = tau_fs / TIME_CONV;
  double cmass0 = (2.0 * tau * tau) * CP_EMAGIC;
  for (int i = 0; i < s->npts; i++) {
    double w = (s->kx[i] == 0) ? 1.0 : 2.0;
    double g2 = s->g2[i];
    cmass[i] = (g2 > ecut_mass) ? (w * g2) * (cmass0 / ecut_mass) : w * cmass0;
  }
}

void oo_sd_update(const oo_sys *s, const double *cmass, double dt, const double *f_, double *psi_) {
  for (int st = 0; st < s->nst; st++) {
    const cplx *f = (const cplx *)f_ + (size_t)st * s->npts;
    cplx *psi = (cplx *)psi_ + (size_t)st * s->npts;
    for (int i = 0; i < s->npts; i++) psi[i] += f[i] * (dt / cmass[i]);
  }
}

double oo_cg_gamma(oo_cg_state *st, double fovlap) {
  /* acceptCgOverlap, CP_State_GSpacePlane.C:131-135 */
  st->fovlap_old = st->fovlap;
  st->fmag_old = st->fmag;
  st->fovlap = fovlap;
  st->fmag = sqrt(fovlap);
  st->iter++;
  /* reset rule, CP_State_GSpacePlane.C:2651-2671 */
  int reset = 0;
  if (st->iter == 1) {
    reset = 1;
    st->num_reset = 0;
  } else if ((st->fmag > 1.05 * st->fmag_old && st->num_reset >= 5) ||
             st->fmag > 2.0 * st->fmag_old) {
    reset = 1;
    st->num_reset = 0;
  }
  st->num_reset++;
  return reset ? 0.0 : st->fovlap / st->fovlap_old;
}

void oo_cg_update(const oo_sys *s, const double *cmass, double dt, double gamma,
                  const double *f_, double *cg_, double *psi_) {
  for (int st = 0; st < s->nst; st++) {
    const cplx *f = (const cplx *)f_ + (size_t)st * s->npts;
    cplx *cg = (cplx *)cg_ + (size_t)st * s->npts;
    cplx *psi = (cplx *)psi_ + (size_t)st * s->npts;
    for (int i = 0; i < s->npts; i++) {
      if (s->g2[i] <= s->ecut_ry) cg[i] = cg[i] * gamma + f[i];
      psi[i] += cg[i] * (dt / cmass[i]);
    }
  }
}

void oo_red_sync(const oo_sys *s, double *psi_) {
  for (int st = 0; st < s->nst; st++) {
    cplx *psi = (cplx *)psi_ + (size_t)st * s->npts;
    for (int i = 0; i < s->npts; i++)
      if (s->partner[i] >= 0) psi[i] = conj(psi[s->partner[i]]);
  }
}

void oo_overlap(const oo_sys *s, const double *psi, double *S) {
  int n = s->nst; size_t K = 2 * (size_t)s->npts;
  double *A = malloc(sizeof(double) * n * K);
  double r2i = 1.0 / sqrt(2.0);
  for (int st = 0; st < n; st++)
    for (size_t k = 0; k < K; k++) {
      double a = psi[(size_t)st * K + k];
      A[(size_t)st * K + k] = (s->kx[k / 2] == 0) ? a * r2i : a;
    }
#pragma omp parallel for schedule(dynamic, 1) num_threads(oo_get_threads())
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      const double *a = A + (size_t)i * K, *b = A + (size_t)j * K;
      double acc = 0.0;
      for (size_t k = 0; k < K; k++) acc += a[k] * b[k];
      S[(size_t)i * n + j] = 2.0 * acc; /* alpha = 2 (flag_dp) */
    }
  free(A);
}

/* DGEMM restatements in column-major flat storage for GENMATMUL */
static void gm_nt(int n, double alpha, const double *X, const double *Y, double beta, double *Z) {
  /* Z = alpha X * Y^T + beta Z */
#pragma omp parallel for schedule(static) num_threads(oo_get_threads())
  for (int c = 0; c < n; c++)
    for (int r = 0; r < n; r++) {
      double acc = 0.0;
      for (int k = 0; k < n; k++) acc += X[r + (size_t)k * n] * Y[c + (size_t)k * n];
      Z[r + (size_t)c * n] = alpha * acc + (beta == 0.0 ? 0.0 : beta * Z[r + (size_t)c * n]);
    }
}

int oo_ns(int n, const double *S, double tol, int max_iter, double *T, double *err) {
  size_t nn = (size_t)n * n;
  double *X = malloc(sizeof(double) * nn), *C = malloc(sizeof(double) * nn), *M3 = malloc(sizeof(double) * nn);
  for (size_t i = 0; i < nn; i++) X[i] = S[i] / 2.0; /* cp_lowdin.C:201 */
  memset(T, 0, sizeof(double) * nn);
  for (int i = 0; i < n; i++) T[i + (size_t)i * n] = 1.0;
  int iter = 0;
  double tol_now = 1.0;
  while (tol_now > tol && iter < max_iter) {
    iter++;
    /* C = 3I - T X^T (cp_lowdin.C:210-217) */
    memset(C, 0, sizeof(double) * nn);
    for (int i = 0; i < n; i++) C[i + (size_t)i * n] = 3.0;
    gm_nt(n, -1.0, T, X, 1.0, C);
    /* X = 0.5 X_old C^T (cp_lowdin.C:220-224) */
    memcpy(M3, X, sizeof(double) * nn);
    gm_nt(n, 0.5, M3, C, 0.0, X);
    /* T = 0.5 C T_old^T (cp_lowdin.C:227-231) */
    memcpy(M3, T, sizeof(double) * nn);
    gm_nt(n, 0.5, C, M3, 0.0, T);
    /* rms change (cp_lowdin.C:234-241; ortho.C:119) */
    double acc = 0.0;
    for (size_t i = 0; i < nn; i++) { double d = M3[i] - T[i]; acc += d * d; }
    tol_now = sqrt(acc / (double)nn);
  }
  free(X); free(C); free(M3);
  if (err) *err = tol_now;
  return (tol_now > tol) ? -iter : iter;
}

void oo_rotate(const oo_sys *s, const double *psi, const double *T, double *out) {
  int n = s->nst; size_t K = 2 * (size_t)s->npts;
#pragma omp parallel for schedule(static) num_threads(oo_get_threads())
  for (long k = 0; k < (long)K; k++)
    for (int j = 0; j < n; j++) {
      double acc = 0.0;
      for (int i = 0; i < n; i++) acc += psi[(size_t)i * K + k] * T[(size_t)i * n + j];
      out[(size_t)j * K + k] = acc;
    }
}

double oo_eloc(const oo_sys *s, const double *v, const double *rho) {
  size_t N = (size_t)s->nx * s->ny * s->nz;
  double acc = 0.0;
  for (size_t r = 0; r < N; r++) acc += v[r] * rho[r];
  return acc * s->vol / (double)N;
}

/* ---------------- V_KS from rho (density chares) ---------------- */

/* cp_hart_ext.C:433-475: row i = clamp(floor((g-gmin)/dg)+1, 1, nspl)
 * (1-based), h = g - ((i-1) dg + gmin), Horner in h */
double oo_vps_eval(const oo_local_pseudo *lp, int sp, double g) {
  int ns = lp->nspline;
  int i = (int)((g - lp->gmin) / lp->dg + 1.0);
  if (i > ns) i = ns;
  if (i < 1) i = 1;
  double h = g - ((double)(i - 1) * lp->dg + lp->gmin);
  const double *c = lp->vps + (size_t)sp * 4 * ns;
  int k = i - 1;
  return ((c[3 * ns + k] * h + c[2 * ns + k]) * h + c[ns + k]) * h + c[k];
}

/* signed frequency of grid index i */
static int sfreq(int i, int n) { return i <= (n - 1) / 2 ? i : i - n; }

void oo_vloc_g(const oo_sys *s, const oo_local_pseudo *lp, double *vloc_) {
  cplx *vloc = (cplx *)vloc_;
  int Nx = s->nx, Ny = s->ny, Nz = s->nz, nh = Nx / 2 + 1;
  double ecut4 = 4.0 * s->ecut_ry, tpi = 2.0 * M_PI;
  const double *hi = s->hmati;
#pragma omp parallel for num_threads(oo_get_threads()) schedule(static)
  for (int z = 0; z < Nz; z++)
    for (int y = 0; y < Ny; y++)
      for (int x = 0; x < nh; x++) {
        int a = x, b = sfreq(y, Ny), c = sfreq(z, Nz);
        double gx = tpi * (a * hi[0] + b * hi[1] + c * hi[2]);
        double gy = tpi * (a * hi[3] + b * hi[4] + c * hi[5]);
        double gz = tpi * (a * hi[6] + b * hi[7] + c * hi[8]);
        double g2 = gx * gx + gy * gy + gz * gz;
        cplx v = 0;
        if (g2 == 0.0) {
          for (int ia = 0; ia < lp->natoms; ia++) v += lp->vps_g0[lp->species[ia]];
        } else if (g2 <= ecut4) {
          double g = sqrt(g2);
          for (int ia = 0; ia < lp->natoms; ia++) {
            const double *R = lp->pos + 3 * ia;
            double arg = gx * R[0] + gy * R[1] + gz * R[2];
            v += oo_vps_eval(lp, lp->species[ia], g) * (cos(arg) - sin(arg) * I);
          }
        }
        vloc[((size_t)z * Ny + y) * nh + x] = v / s->vol;
      }
}

/* ---------------- non-local KB, l = 0 ---------------- */

double oo_nl_proj(const oo_nonlocal *nl, int sp, double g) {
  if (g == 0.0) return nl->proj_g0[sp];
  int ns = nl->nspline;
  int i = (int)((g - nl->gmin) / nl->dg) + 1;
  if (i > ns) i = ns;
  if (i < 1) i = 1;
  double h = g - ((double)(i - 1) * nl->dg + nl->gmin);
  const double *c = nl->proj + (size_t)sp * 4 * ns;
  int k = i - 1;
  return ((c[3 * ns + k] * h + c[2 * ns + k]) * h + c[ns + k]) * h + c[k];
}

/* G of stored point p (cp_struct_fact.C:160-163) */
static void nl_gvec(const oo_sys *s, int p, double G[3]) {
  const double tpi = 2.0 * M_PI, *hi = s->hmati;
  int a = s->kx[p], b = s->ky[p], c = s->kz[p];
  G[0] = tpi * (a * hi[0] + b * hi[1] + c * hi[2]);
  G[1] = tpi * (a * hi[3] + b * hi[4] + c * hi[5]);
  G[2] = tpi * (a * hi[6] + b * hi[7] + c * hi[8]);
}

/* sf[I][p] = P y00 exp(-i G.R_I); gq[p][3] = G */
static void nl_table(const oo_sys *s, const oo_nonlocal *nl, cplx *sf, double *gq) {
  const double y00 = 1.0 / sqrt(4.0 * M_PI);
  int np = s->npts, na = nl->natoms;
#pragma omp parallel for num_threads(oo_get_threads()) schedule(static)
  for (int p = 0; p < np; p++) {
    double G[3];
    nl_gvec(s, p, G);
    gq[3 * (size_t)p] = G[0];
    gq[3 * (size_t)p + 1] = G[1];
    gq[3 * (size_t)p + 2] = G[2];
    double g = sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]);
    for (int ia = 0; ia < na; ia++) {
      const double *R = nl->pos + 3 * (size_t)ia;
      double arg = R[0] * G[0] + R[1] * G[1] + R[2] * G[2];
      double pr = oo_nl_proj(nl, nl->species[ia], g) * y00;
      sf[(size_t)ia * np + p] = pr * (cos(arg) - sin(arg) * I);
    }
  }
}

void oo_nl_zmat(const oo_sys *s, const oo_nonlocal *nl, const double *psi_, double *z,
                double *dz) {
  const cplx *psi = (const cplx *)psi_;
  int np = s->npts, na = nl->natoms, nst = s->nst;
  cplx *sf = malloc(sizeof(cplx) * (size_t)na * np);
  double *gq = malloc(sizeof(double) * 3 * (size_t)np);
  nl_table(s, nl, sf, gq);
#pragma omp parallel for num_threads(oo_get_threads()) schedule(static)
  for (int st = 0; st < nst; st++) {
    const cplx *ps = psi + (size_t)st * np;
    for (int ia = 0; ia < na; ia++) {
      const cplx *f = sf + (size_t)ia * np;
      double acc = 0.0, d[3] = {0.0, 0.0, 0.0};
      for (int p = 0; p < np; p++) {
        double w = s->kx[p] == 0 ? 1.0 : 2.0;
        /* psi conj(SF) = psi P y00 exp(+i G.R) */
        cplx t = ps[p] * conj(f[p]);
        acc += w * creal(t);
        /* d/dR Re(psi P y00 exp(i G.R)) = -G Im(...) */
        for (int q = 0; q < 3; q++) d[q] -= w * gq[3 * (size_t)p + q] * cimag(t);
      }
      z[(size_t)st * na + ia] = acc;
      if (dz)
        for (int q = 0; q < 3; q++) dz[((size_t)st * 3 + q) * na + ia] = d[q];
    }
  }
  free(sf);
  free(gq);
}

double oo_nl_apply(const oo_sys *s, const oo_nonlocal *nl, const double *psi, double *f_,
                   double *fatm) {
  int np = s->npts, na = nl->natoms, nst = s->nst;
  double *z = malloc(sizeof(double) * (size_t)nst * na);
  double *dz = fatm ? malloc(sizeof(double) * (size_t)nst * 3 * na) : NULL;
  oo_nl_zmat(s, nl, psi, z, dz);
  double enl = 0.0;
  for (int st = 0; st < nst; st++) {
    double e = 0.0;
    for (int ia = 0; ia < na; ia++) {
      double norm = nl->vnorm[nl->species[ia]] / s->vol;
      double zz = z[(size_t)st * na + ia];
      e += zz * zz * norm;
      z[(size_t)st * na + ia] = 2.0 * norm * zz; /* Z' */
    }
    enl += e * s->occ[st];
  }
  if (fatm)
    for (int st = 0; st < nst; st++)
      for (int ia = 0; ia < na; ia++)
        for (int q = 0; q < 3; q++)
          fatm[3 * (size_t)ia + q] -=
              z[(size_t)st * na + ia] * dz[((size_t)st * 3 + q) * na + ia];
  if (f_) {
    cplx *f = (cplx *)f_;
    cplx *sf = malloc(sizeof(cplx) * (size_t)na * np);
    double *gq = malloc(sizeof(double) * 3 * (size_t)np);
    nl_table(s, nl, sf, gq);
#pragma omp parallel for num_threads(oo_get_threads()) schedule(static)
    for (int st = 0; st < nst; st++)
      for (int p = 0; p < np; p++) {
        double w = s->kx[p] == 0 ? 1.0 : 2.0;
        cplx acc = 0.0;
        for (int ia = 0; ia < na; ia++) acc += z[(size_t)st * na + ia] * sf[(size_t)ia * np + p];
        f[(size_t)st * np + p] -= w * acc;
      }
    free(sf);
    free(gq);
  }
  free(z);
  free(dz);
  return enl;
}

/* cp_pz_lda.C:101-153, per point */
void oo_pz_lda(size_t n, const double *rho, double *vxc, double sums[3]) {
  const double gamma = -0.14230, beta1 = 1.05290, beta2 = 0.33340;
  const double au = 0.03110, bu = -0.0480, cu = 0.00200, du = -0.01160;
  const double power = 1.0 / 3.0, fpi = 4.0 * M_PI;
  const double rat1 = (7.0 / 6.0) * beta1, rat2 = (4.0 / 3.0) * beta2;
  double ex = 0.0, ec = 0.0, vx = 0.0;
  for (size_t i = 0; i < n; i++) {
    double r = rho[i];
    double rs = pow(3.0 / (fpi * r), power);
    double sq = sqrt(rs);
    double xf = -pow(3.0 * r / M_PI, power);
    double cf, mf;
    if (rs >= 1.0) {
      double c1 = 1.0 + beta1 * sq + beta2 * rs;
      double c2 = 1.0 + rat1 * sq + rat2 * rs;
      cf = gamma / c1;
      mf = cf * (c2 / c1);
    } else {
      double l = log(rs);
      cf = au * l + bu + cu * rs * l + du * rs;
      mf = au * l + (bu - power * au) + 2.0 * power * cu * rs * l + power * (2.0 * du - cu) * rs;
    }
    ex += xf * r;
    ec += cf * r;
    vxc[i] = xf + mf;
    vx += (xf + mf) * r;
  }
  sums[0] = ex; sums[1] = ec; sums[2] = vx;
}

void oo_vks(const oo_sys *s, const double *rho, const double *vloc_, double *v,
            oo_vks_energies *e) {
  const cplx *vloc = (const cplx *)vloc_;
  int Nx = s->nx, Ny = s->ny, Nz = s->nz, nh = Nx / 2 + 1;
  size_t N = (size_t)Nx * Ny * Nz, NH = (size_t)Nz * Ny * nh;
  double sums[3];
  oo_pz_lda(N, rho, v, sums);  /* v = V_xc */
  cplx *h = malloc(sizeof(cplx) * NH);
  int nmax = Nx > Ny ? Nx : Ny; if (Nz > nmax) nmax = Nz;
  const double sc = s->vol / (double)N;
  /* rho_G: x (complex FFT of the real row), y, z; FFTW_FORWARD sign */
#pragma omp parallel num_threads(oo_get_threads())
  {
    cplx *row = malloc(sizeof(cplx) * nmax), *w = malloc(sizeof(cplx) * 3 * nmax);
#pragma omp for schedule(static)
    for (int zy = 0; zy < Nz * Ny; zy++) {
      for (int x = 0; x < Nx; x++) row[x] = rho[(size_t)zy * Nx + x] * sc;
      fft_exec(&s->px, -1, row, w);
      for (int x = 0; x < nh; x++) h[(size_t)zy * nh + x] = row[x];
    }
#pragma omp for schedule(static)
    for (int zx = 0; zx < Nz * nh; zx++) {
      int z = zx / nh, x = zx % nh;
      for (int y = 0; y < Ny; y++) row[y] = h[((size_t)z * Ny + y) * nh + x];
      fft_exec(&s->py, -1, row, w);
      for (int y = 0; y < Ny; y++) h[((size_t)z * Ny + y) * nh + x] = row[y];
    }
#pragma omp for schedule(static)
    for (int yx = 0; yx < Ny * nh; yx++) {
      for (int z = 0; z < Nz; z++) row[z] = h[(size_t)z * Ny * nh + yx];
      fft_exec(&s->pz, -1, row, w);
      for (int z = 0; z < Nz; z++) h[(size_t)z * Ny * nh + yx] = row[z];
    }
    free(row); free(w);
  }
  /* Hartree + local pseudo in G (cp_hart_ext.C:186-330) */
  double ehart = 0.0, eext = 0.0, ecut4 = 4.0 * s->ecut_ry, fpi = 4.0 * M_PI;
  for (int z = 0; z < Nz; z++)
    for (int y = 0; y < Ny; y++)
      for (int x = 0; x < nh; x++) {
        size_t i = ((size_t)z * Ny + y) * nh + x;
        double g2 = full_g2(s->hmati, x, sfreq(y, Ny), sfreq(z, Nz));
        cplx r = h[i], vl = vloc ? vloc[i] : 0;
        double w = x == 0 ? 1.0 : 2.0;
        if (g2 == 0.0) {
          h[i] = creal(vl);
          eext += creal(vl) * creal(r);
        } else if (g2 <= ecut4) {
          double hf = fpi / (g2 * s->vol);
          ehart += hf * (creal(r) * creal(r) + cimag(r) * cimag(r)) * w;
          eext += creal(r * conj(vl)) * w;
          h[i] = r * hf + vl;
        } else {
          h[i] = 0;
        }
      }
  /* back: z, y (FFTW_BACKWARD), Hermitian x c2r (Im of kx = 0 and of the
   * Nyquist column dropped), added to V_xc */
#pragma omp parallel num_threads(oo_get_threads())
  {
    cplx *row = malloc(sizeof(cplx) * nmax), *w = malloc(sizeof(cplx) * 3 * nmax);
#pragma omp for schedule(static)
    for (int yx = 0; yx < Ny * nh; yx++) {
      for (int z = 0; z < Nz; z++) row[z] = h[(size_t)z * Ny * nh + yx];
      fft_exec(&s->pz, +1, row, w);
      for (int z = 0; z < Nz; z++) h[(size_t)z * Ny * nh + yx] = row[z];
    }
#pragma omp for schedule(static)
    for (int zx = 0; zx < Nz * nh; zx++) {
      int z = zx / nh, x = zx % nh;
      for (int y = 0; y < Ny; y++) row[y] = h[((size_t)z * Ny + y) * nh + x];
      fft_exec(&s->py, +1, row, w);
      for (int y = 0; y < Ny; y++) h[((size_t)z * Ny + y) * nh + x] = row[y];
    }
#pragma omp for schedule(static)
    for (int zy = 0; zy < Nz * Ny; zy++) {
      const cplx *hr = h + (size_t)zy * nh;
      row[0] = creal(hr[0]);
      for (int x = 1; x < nh; x++) row[x] = hr[x];
      if (Nx % 2 == 0) row[Nx / 2] = creal(hr[Nx / 2]);
      for (int x = nh; x < Nx; x++) row[x] = conj(hr[Nx - x]);
      fft_exec(&s->px, +1, row, w);
      for (int x = 0; x < Nx; x++) v[(size_t)zy * Nx + x] += creal(row[x]);
    }
    free(row); free(w);
  }
  free(h);
  if (e) {
    e->ehart = 0.5 * ehart;
    e->eext = eext;
    e->exc = (0.75 * sums[0] + sums[1]) * sc;
    e->muxc = sums[2] * sc;
  }
}

int oo_step(const oo_sys *s, const oo_step_opts *o, const double *v, double *psi,
            double *rho_o, double *S_o, double *T_o, double *f_o, oo_step_out *out) {
  size_t N = (size_t)s->nx * s->ny * s->nz;
  int n = s->nst; size_t K = 2 * (size_t)s->npts;
  double *psir = malloc(sizeof(double) * N * n);
  double *rho = rho_o ? rho_o : malloc(sizeof(double) * N);
  double *f = f_o ? f_o : malloc(sizeof(double) * K * n);
  double *S = S_o ? S_o : malloc(sizeof(double) * n * n);
  double *T = T_o ? T_o : malloc(sizeof(double) * n * n);
  double *cm = malloc(sizeof(double) * s->npts);
  oo_g2r_all(s, psi, psir);
  oo_density(s, psir, rho);
  oo_vks_energies ve = {0, 0, 0, 0};
  double *vk = NULL;
  if (o->vks_from_rho) {
    vk = o->vks_out ? o->vks_out : malloc(sizeof(double) * N);
    oo_vks(s, rho, o->vloc, vk, &ve);
    v = vk;
  }
  oo_vpsi_r2g_all(s, psir, v, f);
  double eke = oo_eke_force(s, psi, f);
  double eloc = o->vks_from_rho ? ve.eext : oo_eloc(s, v, rho);
  double enl = o->nl ? oo_nl_apply(s, o->nl, psi, f, NULL) : 0.0;
  if (vk && !o->vks_out) free(vk);
  if (o->use_lambda) {
    double *lam = malloc(sizeof(double) * n * n);
    oo_lambda(s, psi, f, lam);
    oo_lambda_apply(s, psi, lam, f);
    free(lam);
  }
  double fn = 0.0;
  for (size_t i = 0; i < K * n; i++) fn += f[i] * f[i];
  oo_cmass(s, o->tau_fs, o->ecut_mass, cm);
  if (o->use_cg && o->cg_state && o->cg) {
    const double gamma = oo_cg_gamma(o->cg_state, fn);
    oo_cg_update(s, cm, o->dt, gamma, f, o->cg, psi);
  } else {
    oo_sd_update(s, cm, o->dt, f, psi);
  }
  oo_red_sync(s, psi);
  oo_overlap(s, psi, S);
  double err = 0.0;
  int it = oo_ns(n, S, o->ns_tol, o->ns_max_iter, T, &err);
  double *tmp = malloc(sizeof(double) * K * n);
  oo_rotate(s, psi, T, tmp);
  memcpy(psi, tmp, sizeof(double) * K * n);
  free(tmp);
  if (out) {
    out->eke = eke; out->eloc = eloc; out->eelec = eke + eloc + ve.ehart + ve.exc + enl;
    out->ehart = ve.ehart; out->exc = ve.exc; out->enl = enl;
    out->fnorm = sqrt(fn); out->ns_iters = it < 0 ? -it : it; out->ns_err = err;
  }
  free(psir); free(cm);
  if (!rho_o) free(rho);
  if (!f_o) free(f);
  if (!S_o) free(S);
  if (!T_o) free(T);
  return it < 0 ? 1 : 0;
}
