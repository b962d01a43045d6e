// oa_vks.hip -- V_KS(r) from rho(r) on the device: the work of the
// reference's density chares for a fixed-ion Gamma-point step.
//
//   k_vks_x<-1> : per z-plane, pairs of rows: PZ-LDA V_xc(r) into v and its
//                 energy sums (CP_exc_calc, cp_pz_lda.C:101-167); rho Omega/N
//                 (CP_Rho_RealSpacePlane.C:850), x r2c -> hx[z][kx][y]
//   k_vks_y<-1> : y FFT of the kx columns -> hl[z][kx][ky] (density box)
//   k_vks_lines : z FFT of each (kx,ky) line, rho(G) -> 4 pi rho/(G^2 Omega)
//                 + V_loc(G) on |G|^2 <= 4 ecut (CP_hart_eext_calc,
//                 cp_hart_ext.C:186-330) with E_H and E_ext sums, z FFT back
//   k_vks_y<+1> : y FFT back -> hx
//   k_vks_xb    : Hermitian x c2r of row pairs, added to V_xc
//                 (acceptHartVks, CP_Rho_RealSpacePlane.C:1365-1390)
//   k_vks_sum   : fixed-order reduction of the partial sums into StepDev
//
// One grid per step (not one per state), so the passes are plain batched
// Stockham transforms on LDS: each moves ~2 x 8 MB at 100^3, well under 5% of
// the state chain.  All arithmetic is FP64 in both precision modes.
#include <algorithm>

#include "oa_common.h"
#include "oa_kernels.h"
#include "oa_stockham.h"

namespace oa {
namespace {

constexpr int kVT = 256;
constexpr int kVW = kVT / 64;
// Staging loops run KU iterations per round with every load issued first,
// from a clamped (valid) index, and the mask applied afterwards: a load in
// a divergent branch is waited for on its own.
constexpr int KU = 4;

__device__ __forceinline__ int sfreq(int i, int n) { return i <= (n - 1) / 2 ? i : i - n; }

__device__ __forceinline__ double gsq(const double* hi, int a, int b, int c) {
  const double tpi = 6.283185307179586476925;
  const double xk = (a * hi[0] + b * hi[1] + c * hi[2]) * tpi;
  const double yk = (a * hi[3] + b * hi[4] + c * hi[5]) * tpi;
  const double zk = (a * hi[6] + b * hi[7] + c * hi[8]) * tpi;
  return xk * xk + yk * yk + zk * zk;
}

// Perdew-Zunger 1981 LDA per point: exchange factor xf, correlation energy
// factor cf and potential factor mf (cp_pz_lda.C:48-54, 114-138).  One cube
// root serves both powers: rs = (3/(4 pi))^(1/3) / rho^(1/3) and
// xf = -(3/pi)^(1/3) rho^(1/3).
__device__ __forceinline__ void pz_lda(double r, double& xf, double& cf, double& mf) {
  const double gamma = -0.14230, beta1 = 1.05290, beta2 = 0.33340;
  const double au = 0.03110, bu = -0.0480, cu = 0.00200, du = -0.01160;
  const double power = 1.0 / 3.0;
  const double rat1 = (7.0 / 6.0) * beta1, rat2 = (4.0 / 3.0) * beta2;
  const double c_rs = 0.62035049089940001667;  // (3/(4 pi))^(1/3)
  const double c_x = 0.98474502184269654118;   // (3/pi)^(1/3)
  const double cr = cbrt(r);
  const double rs = c_rs / cr;
  const double sq = sqrt(rs);
  xf = -c_x * cr;
  if (rs >= 1.0) {
    const double c1 = 1.0 + beta1 * sq + beta2 * rs;
    const double c2 = 1.0 + rat1 * sq + rat2 * rs;
    cf = gamma / c1;
    mf = cf * (c2 / c1);
  } else {
    const double l = log(rs);
    cf = au * l + bu + cu * rs * l + du * rs;
    mf = au * l + (bu - power * au) + 2.0 * power * cu * rs * l + power * (2.0 * du - cu) * rs;
  }
}

// sum of NV values over the workgroup (fixed order); result valid on thread 0
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* sh) {
#pragma unroll
  for (int k = 0; k < NV; k++)
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) sh[w * NV + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double acc = 0.0;
      for (int i = 0; i < kVW; i++) acc += sh[i * NV + k];
      v[k] = acc;
    }
}

// x pass forward: grid (nz, ceil(npair / pr)).  LDS: tw[nx], A[pr][nx], B[pr][nx].
__global__ __launch_bounds__(kVT) void k_vks_x(VksArgs a) {
  extern __shared__ double2 vks_smem[];
  cpx<double>* tw = reinterpret_cast<cpx<double>*>(vks_smem);
  cpx<double>* A = tw + a.nx;
  cpx<double>* B = A + a.pr * a.nx;
  __shared__ double red[kVW * 3];
  const int z = blockIdx.x, p0 = blockIdx.y * a.pr;
  const int np = min(a.pr, (a.ny + 1) / 2 - p0);
  load_tw<double, kVT>(tw, a.twx, a.nx);
  const double sc = a.vol / ((double)a.nx * a.ny * a.nz);
  double acc[3] = {0.0, 0.0, 0.0};
  const int cnt = a.pr * a.nx;
  for (int i0 = threadIdx.x; i0 < cnt; i0 += KU * kVT) {
    double rv[KU][2];
    size_t iv[KU][2];
    bool ok[KU][2];
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int i = i0 + u * kVT, q = i / a.nx, x = i - q * a.nx;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int y = 2 * (p0 + q) + h;
        ok[u][h] = i < cnt && q < np && y < a.ny;
        iv[u][h] = ok[u][h] ? ((size_t)z * a.ny + y) * a.nx + x : 0;
        rv[u][h] = a.rho[iv[u][h]];
      }
    }
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int i = i0 + u * kVT;
      if (i >= cnt) break;
      cpx<double> val = {0.0, 0.0};
#pragma unroll
      for (int h = 0; h < 2; h++) {
        if (ok[u][h]) {
          const double r = rv[u][h];
          double xf, cf, mf;
          pz_lda(r, xf, cf, mf);
          a.v[iv[u][h]] = xf + mf;
          acc[0] += xf * r;
          acc[1] += cf * r;
          acc[2] += (xf + mf) * r;
          if (h == 0) val.x = r * sc; else val.y = r * sc;
        }
      }
      A[i] = val;
    }
  }
  __syncthreads();
  const cpx<double>* R = fft_lds<double, -1, kVT>(A, B, a.px, a.pr, a.nx, tw);
  // Z = X_a + i X_b: X_a(k) = (Z_k + conj Z_-k)/2, X_b(k) = (Z_k - conj Z_-k)/(2i)
  const int rows = 2 * a.pr;
  for (int i = threadIdx.x; i < a.nkx * rows; i += kVT) {
    const int k = i / rows, yy = i - k * rows;
    const int q = yy >> 1, y = 2 * p0 + yy;
    if (q < np && y < a.ny) {
      const cpx<double> zk = R[q * a.nx + k];
      const cpx<double> zm = R[q * a.nx + (k == 0 ? 0 : a.nx - k)];
      double2 o;
      if ((yy & 1) == 0)
        o = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
      else
        o = make_double2(0.5 * (zk.y + zm.y), -0.5 * (zk.x - zm.x));
      a.hx[((size_t)z * a.nkx + k) * a.ny + y] = o;
    }
  }
  block_sum<3>(acc, red);
  if (threadIdx.x == 0) {
    double* p = a.part_x + (size_t)(blockIdx.x * gridDim.y + blockIdx.y) * 3;
    p[0] = acc[0];
    p[1] = acc[1];
    p[2] = acc[2];
  }
}

// y pass: grid (nz, ceil(nkx / cb)).  S = -1: hx -> hl (ky box only);
// S = +1: hl (zero outside the box) -> hx.  LDS: tw[ny], A[cb][ny], B[cb][ny].
template <int S>
__global__ __launch_bounds__(kVT) void k_vks_y(VksArgs a) {
  extern __shared__ double2 vks_smem[];
  cpx<double>* tw = reinterpret_cast<cpx<double>*>(vks_smem);
  cpx<double>* A = tw + a.ny;
  cpx<double>* B = A + a.cb * a.ny;
  const int z = blockIdx.x, k0 = blockIdx.y * a.cb;
  const int nc = min(a.cb, a.nkx - k0);
  load_tw<double, kVT>(tw, a.twy, a.ny);
  const int cnt = a.cb * a.ny;
  for (int i0 = threadIdx.x; i0 < cnt; i0 += KU * kVT) {
    double2 v[KU];
    bool ok[KU];
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int i = i0 + u * kVT, c = i / a.ny, y = i - c * a.ny;
      const size_t col = (size_t)z * a.nkx + k0 + c;
      size_t idx;
      if (S < 0) {
        ok[u] = i < cnt && c < nc;
        idx = col * a.ny + y;
      } else {
        const int ky = sfreq(y, a.ny);
        ok[u] = i < cnt && c < nc && ky >= -a.kyr && ky <= a.kyr;
        idx = col * a.nky + ky + a.kyr;
      }
      v[u] = (S < 0 ? a.hx : a.hl)[ok[u] ? idx : 0];
    }
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int i = i0 + u * kVT;
      if (i < cnt) A[i] = ok[u] ? cpx<double>{v[u].x, v[u].y} : cpx<double>{0.0, 0.0};
    }
  }
  __syncthreads();
  const cpx<double>* R = fft_lds<double, S, kVT>(A, B, a.py, a.cb, a.ny, tw);
  if (S < 0) {
    for (int i = threadIdx.x; i < a.cb * a.nky; i += kVT) {
      const int c = i / a.nky, j = i - c * a.nky;
      if (c < nc) {
        const int ky = j - a.kyr;
        const int y = ky < 0 ? ky + a.ny : ky;
        const cpx<double> r = R[c * a.ny + y];
        a.hl[((size_t)z * a.nkx + k0 + c) * a.nky + j] = make_double2(r.x, r.y);
      }
    }
  } else {
    for (int i = threadIdx.x; i < a.cb * a.ny; i += kVT) {
      const int c = i / a.ny, y = i - c * a.ny;
      if (c < nc) {
        const cpx<double> r = R[i];
        a.hx[((size_t)z * a.nkx + k0 + c) * a.ny + y] = make_double2(r.x, r.y);
      }
    }
  }
}

// z lines: grid (nkx, ceil(nky / lb)).  LDS: tw[nz], A[lb][nz], B[lb][nz].
__global__ __launch_bounds__(kVT) void k_vks_lines(VksArgs a) {
  extern __shared__ double2 vks_smem[];
  cpx<double>* tw = reinterpret_cast<cpx<double>*>(vks_smem);
  cpx<double>* A = tw + a.nz;
  cpx<double>* B = A + a.lb * a.nz;
  __shared__ double red[kVW * 2];
  const int kx = blockIdx.x, j0 = blockIdx.y * a.lb;
  const int nl = min(a.lb, a.nky - j0);
  load_tw<double, kVT>(tw, a.twz, a.nz);
  const int cnt = a.lb * a.nz;
  for (int i0 = threadIdx.x; i0 < cnt; i0 += KU * kVT) {
    double2 v[KU];
    bool ok[KU];
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int i = i0 + u * kVT, z = i / a.lb, l = i - z * a.lb;
      ok[u] = i < cnt && l < nl;
      v[u] = a.hl[ok[u] ? ((size_t)z * a.nkx + kx) * a.nky + j0 + l : 0];
    }
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int i = i0 + u * kVT, z = i / a.lb, l = i - z * a.lb;
      if (i < cnt) A[l * a.nz + z] = ok[u] ? cpx<double>{v[u].x, v[u].y} : cpx<double>{0.0, 0.0};
    }
  }
  __syncthreads();
  cpx<double>* R = fft_lds<double, -1, kVT>(A, B, a.pz, a.lb, a.nz, tw);
  cpx<double>* O = (R == A) ? B : A;
  const double fpi = 4.0 * 3.14159265358979323846;
  const double w = kx == 0 ? 1.0 : 2.0;
  double acc[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < cnt; i += kVT) {
    const int l = i / a.nz, z = i - l * a.nz;
    // V_loc loaded before the branch (clamped row): not waited for alone
    const double2 vl = a.vloc[((size_t)kx * a.nky + j0 + min(l, nl - 1)) * a.nz + z];
    cpx<double> out = {0.0, 0.0};
    if (l < nl) {
      const int ky = j0 + l - a.kyr, kz = sfreq(z, a.nz);
      const double g2 = gsq(a.hmati, kx, ky, kz);
      const cpx<double> r = R[i];
      if (g2 == 0.0) {
        out = {vl.x, 0.0};
        acc[1] += vl.x * r.x;
      } else if (g2 <= a.ecut4) {
        const double hf = fpi / (g2 * a.vol);
        acc[0] += hf * (r.x * r.x + r.y * r.y) * w;
        acc[1] += (r.x * vl.x + r.y * vl.y) * w;  // Re(rho conj V_loc)
        out = {r.x * hf + vl.x, r.y * hf + vl.y};
      }
    }
    O[i] = out;
  }
  __syncthreads();
  const cpx<double>* R2 = fft_lds<double, 1, kVT>(O, R, a.pz, a.lb, a.nz, tw);
  for (int i = threadIdx.x; i < a.lb * a.nz; i += kVT) {
    const int z = i / a.lb, l = i - z * a.lb;
    if (l < nl) {
      const cpx<double> r = R2[l * a.nz + z];
      a.hl[((size_t)z * a.nkx + kx) * a.nky + j0 + l] = make_double2(r.x, r.y);
    }
  }
  block_sum<2>(acc, red);
  if (threadIdx.x == 0) {
    double* p = a.part_l + (size_t)(blockIdx.x * gridDim.y + blockIdx.y) * 2;
    p[0] = acc[0];
    p[1] = acc[1];
  }
}

// x pass back (FFTW c2r: Im of kx = 0 dropped): grid (nz, ceil(npair / pr)).
// LDS: tw[nx], A[pr][nx], B[max(pr nx, nkx 2 pr)]; the raw half rows are
// staged in B.
__global__ __launch_bounds__(kVT) void k_vks_xb(VksArgs a) {
  extern __shared__ double2 vks_smem[];
  cpx<double>* tw = reinterpret_cast<cpx<double>*>(vks_smem);
  cpx<double>* A = tw + a.nx;
  cpx<double>* B = A + a.pr * a.nx;
  const int z = blockIdx.x, p0 = blockIdx.y * a.pr;
  const int np = min(a.pr, (a.ny + 1) / 2 - p0);
  const int rows = 2 * a.pr;
  load_tw<double, kVT>(tw, a.twx, a.nx);
  // B[k][yy] = hx[z][k][2 p0 + yy] (coalesced along y)
  const int cnt = a.nkx * rows;
  for (int i0 = threadIdx.x; i0 < cnt; i0 += KU * kVT) {
    double2 v[KU];
    bool ok[KU];
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int i = i0 + u * kVT, k = i / rows, yy = i - k * rows, y = 2 * p0 + yy;
      ok[u] = i < cnt && (yy >> 1) < np && y < a.ny;
      v[u] = a.hx[ok[u] ? ((size_t)z * a.nkx + k) * a.ny + y : 0];
    }
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int i = i0 + u * kVT;
      if (i < cnt) B[i] = ok[u] ? cpx<double>{v[u].x, v[u].y} : cpx<double>{0.0, 0.0};
    }
  }
  __syncthreads();
  // Z_k = X_a(k) + i X_b(k), Z_-k = conj X_a(k) + i conj X_b(k)
  for (int i = threadIdx.x; i < a.pr * a.nx; i += kVT) {
    const int q = i / a.nx, x = i - q * a.nx;
    const int k = x <= a.nx / 2 ? x : x - a.nx;
    const int m = k < 0 ? -k : k;
    cpx<double> zv = {0.0, 0.0};
    if (m < a.nkx) {
      cpx<double> xa = B[m * rows + 2 * q], xb = B[m * rows + 2 * q + 1];
      if (k == 0) {
        xa.y = 0.0;
        xb.y = 0.0;
      } else if (k < 0) {
        xa.y = -xa.y;
        xb.y = -xb.y;
      }
      zv = {xa.x - xb.y, xa.y + xb.x};
    }
    A[i] = zv;
  }
  __syncthreads();
  const cpx<double>* R = fft_lds<double, 1, kVT>(A, B, a.px, a.pr, a.nx, tw);
  const int cnt2 = a.pr * a.nx;
  for (int i0 = threadIdx.x; i0 < cnt2; i0 += KU * kVT) {
    double v0[KU], v1[KU];
    size_t i0v[KU], i1v[KU];
#pragma unroll
    for (int u = 0; u < KU; u++) {  // V_xc of both rows, loads first
      const int i = i0 + u * kVT, q = i / a.nx, x = i - q * a.nx, y = 2 * (p0 + q);
      const bool ok = i < cnt2 && q < np;
      i0v[u] = ok ? ((size_t)z * a.ny + y) * a.nx + x : 0;
      i1v[u] = (ok && y + 1 < a.ny) ? ((size_t)z * a.ny + y + 1) * a.nx + x : 0;
      v0[u] = a.v[i0v[u]];
      v1[u] = a.v[i1v[u]];
    }
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int i = i0 + u * kVT, q = i / a.nx, y = 2 * (p0 + q);
      if (i < cnt2 && q < np) {
        const cpx<double> r = R[i];
        a.v[i0v[u]] = v0[u] + r.x;
        if (y + 1 < a.ny) a.v[i1v[u]] = v1[u] + r.y;
      }
    }
  }
}

// fixed-order sums of the partials: E_xc = (3/4 ex + ec) Omega/N,
// muxc = sum V_xc rho Omega/N (cp_pz_lda.C:158-167), E_H = half the sum,
// E_ext (cp_hart_ext.C:385-387)
__global__ __launch_bounds__(kVT) void k_vks_sum(const double* __restrict__ px, int nbx,
                                                  const double* __restrict__ pl, int nbl,
                                                  double sc, StepDev* res) {
  __shared__ double red[kVW * 5];
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < nbx; i += kVT) {
    acc[0] += px[3 * i];
    acc[1] += px[3 * i + 1];
    acc[2] += px[3 * i + 2];
  }
  for (int i = threadIdx.x; i < nbl; i += kVT) {
    acc[3] += pl[2 * i];
    acc[4] += pl[2 * i + 1];
  }
  block_sum<5>(acc, red);
  if (threadIdx.x == 0) {
    res->exc = (0.75 * acc[0] + acc[1]) * sc;
    res->muxc = acc[2] * sc;
    res->ehart = 0.5 * acc[3];
    res->eloc = acc[4];
  }
}

// spline of one species at |G| = g (CP_get_vpsnow, cp_hart_ext.C:433-475)
__device__ __forceinline__ double vps_eval(const PseudoDev& p, int sp, double g) {
  const int ns = p.nspline;
  int i = (int)((g - p.gmin) / p.dg + 1.0);
  i = min(max(i, 1), ns);
  const double h = g - ((double)(i - 1) * p.dg + p.gmin);
  const double* c = p.vps + (size_t)sp * 4 * ns;
  const int k = i - 1;
  return ((c[3 * ns + k] * h + c[2 * ns + k]) * h + c[ns + k]) * h + c[k];
}

// V_loc(G) = (1/Omega) sum_I v_I(|G|) exp(-i G.R_I) on the line box [nkx][nky][nz]
__global__ __launch_bounds__(kVT) void k_vloc(VksArgs a, PseudoDev p, double2* vloc) {
  const size_t n = (size_t)a.nkx * a.nky * a.nz;
  const double tpi = 6.283185307179586476925;
  for (size_t i = (size_t)blockIdx.x * kVT + threadIdx.x; i < n; i += (size_t)gridDim.x * kVT) {
    const int z = (int)(i % a.nz);
    const size_t t = i / a.nz;
    const int j = (int)(t % a.nky), kx = (int)(t / a.nky);
    const int ky = j - a.kyr, kz = sfreq(z, a.nz);
    const double* hi = a.hmati;
    const double gx = tpi * (kx * hi[0] + ky * hi[1] + kz * hi[2]);
    const double gy = tpi * (kx * hi[3] + ky * hi[4] + kz * hi[5]);
    const double gz = tpi * (kx * hi[6] + ky * hi[7] + kz * hi[8]);
    const double g2 = gx * gx + gy * gy + gz * gz;
    double re = 0.0, im = 0.0;
    if (g2 == 0.0) {
      for (int ia = 0; ia < p.natoms; ia++) re += p.vps_g0[p.species[ia]];
    } else if (g2 <= a.ecut4) {
      const double g = sqrt(g2);
      for (int ia = 0; ia < p.natoms; ia++) {
        const double* R = p.pos + 3 * ia;
        const double arg = gx * R[0] + gy * R[1] + gz * R[2];
        double s, c;
        sincos(arg, &s, &c);
        const double f = vps_eval(p, p.species[ia], g);
        re += f * c;
        im -= f * s;
      }
    }
    vloc[i] = make_double2(re / a.vol, im / a.vol);
  }
}

}  // namespace

int vks_nbx(const VksArgs& a) { return a.nz * (((a.ny + 1) / 2 + a.pr - 1) / a.pr); }
int vks_nbl(const VksArgs& a) { return a.nkx * ((a.nky + a.lb - 1) / a.lb); }

// x passes: tw, A and B (B also stages nkx x 2 pr half-row values)
size_t vks_smem_x(const VksArgs& a) {
  return sizeof(double2) * ((size_t)a.nx * (1 + a.pr) +
                            std::max((size_t)a.pr * a.nx, (size_t)2 * a.pr * a.nkx));
}

size_t vks_smem(const VksArgs& a) {
  const size_t sx = vks_smem_x(a) / sizeof(double2);
  const size_t sy = (size_t)a.ny * (1 + 2 * a.cb);
  const size_t sz = (size_t)a.nz * (1 + 2 * a.lb);
  return sizeof(double2) * std::max(sx, std::max(sy, sz));
}

hipError_t vks_build(const VksArgs& a, StepDev* res, hipStream_t st) {
  const int npb = ((a.ny + 1) / 2 + a.pr - 1) / a.pr;
  const size_t shx = vks_smem_x(a);
  const size_t shy = sizeof(double2) * (size_t)a.ny * (1 + 2 * a.cb);
  const size_t shz = sizeof(double2) * (size_t)a.nz * (1 + 2 * a.lb);
  const dim3 gx(a.nz, npb), gy(a.nz, (a.nkx + a.cb - 1) / a.cb),
      gl(a.nkx, (a.nky + a.lb - 1) / a.lb);
  hipLaunchKernelGGL(k_vks_x, gx, dim3(kVT), shx, st, a);
  hipLaunchKernelGGL(k_vks_y<-1>, gy, dim3(kVT), shy, st, a);
  hipLaunchKernelGGL(k_vks_lines, gl, dim3(kVT), shz, st, a);
  hipLaunchKernelGGL(k_vks_y<1>, gy, dim3(kVT), shy, st, a);
  hipLaunchKernelGGL(k_vks_xb, gx, dim3(kVT), shx, st, a);
  const double sc = a.vol / ((double)a.nx * a.ny * a.nz);
  hipLaunchKernelGGL(k_vks_sum, dim3(1), dim3(kVT), 0, st, a.part_x, vks_nbx(a), a.part_l,
                     vks_nbl(a), sc, res);
  return hipGetLastError();
}

hipError_t vloc_build(const VksArgs& a, const PseudoDev& p, double2* vloc, hipStream_t st) {
  const size_t n = (size_t)a.nkx * a.nky * a.nz;
  const int nb = (int)std::min<size_t>((n + kVT - 1) / kVT, 4096);
  hipLaunchKernelGGL(k_vloc, dim3(nb), dim3(kVT), 0, st, a, p, vloc);
  return hipGetLastError();
}

}  // namespace oa
