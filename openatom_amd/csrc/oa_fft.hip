// oa_fft.hip -- state FFT chain of the CP electronic step on CDNA4.
//
// The reference runs a 3-pass pruned FFT per state on the CPU
// (src_charm_driver/fft_slab_ctrl/fftCache.C:460-553, 1099-1319): expand the
// G-sphere into z-lines, z-FFT, transpose into half planes, y-FFT on the
// nplane_x populated columns, conjugate, x c2r; and the mirror image for the
// V(r)psi(r) force.  Here each pass is one batched launch over all states:
//
//   k_zfwd  : sphere gather + z-FFT            psi(G)  -> zbuf[s][z][kx][ky]
//   k_y<-1> : y-FFT, transposing write          zbuf    -> xbuf[s][z][y][kx]
//   k_xc2r  : hermitian pair-row x-FFT, rho     xbuf    -> psiR[s][z][y][x], rho
//   k_xr2c  : V psi product, pair-row x-FFT     psiR,V  -> xbuf
//   k_y<+1> : y-FFT (backward), transposing     xbuf    -> zbuf
//   k_zbwd  : z-FFT, pack, -2/N, kinetic force  zbuf    -> F(G), EKE partials
//
// All 1D transforms are Stockham autosort passes (radix 2/3/4/5) on
// ping-pong LDS buffers, batched over many lines/columns/rows per workgroup.
// The two x passes use the Gamma-point trick of carrying two real rows in
// one complex transform.  rho is accumulated in registers by workgroups that
// own a block of rows and loop over all states, so it is written once and
// deterministically.
#include <algorithm>

#include "oa_common.h"
#include "oa_kernels.h"
#include "oa_stockham.h"

namespace oa {

template <typename T>
__device__ __forceinline__ int zdecode_line(int code) { return (code >> 12) & 0x7ffff; }

// ---------------------------------------------------------------- z passes

// expandGSpace + z FFT (FFTW_FORWARD): fftCache.C:460-507, 1133-1160, then the
// G->R transpose of CP_State_GSpacePlane.C:1449-1518 folded into the store.
template <typename T>
__global__ __launch_bounds__(OA_NTHR) void k_zfwd(Geom g, FftPlan pz, int lpb,
                                                  const double2* __restrict__ psi,
                                                  const int* __restrict__ blk_pt0,
                                                  const int* __restrict__ pt_code,
                                                  const int* __restrict__ line_dst,
                                                  const double2* __restrict__ twz,
                                                  cpx<T>* __restrict__ zbuf) {
  extern __shared__ __align__(16) char smem_raw[];
  cpx<T>* tw = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* a = tw + g.nz;
  cpx<T>* b = a + lpb * g.nz;
  const int s = blockIdx.y, blk = blockIdx.x;
  const int l0 = blk * lpb;
  const int nl = min(lpb, g.nlines - l0);
  load_tw(tw, twz, g.nz);
  for (int i = threadIdx.x; i < nl * g.nz; i += blockDim.x) a[i] = {(T)0, (T)0};
  __syncthreads();
  const int p0 = blk_pt0[blk], p1 = blk_pt0[blk + 1];
  const double2* ps = psi + (size_t)s * g.ld;
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const int code = pt_code[p];
    const double2 v = ps[p];
    a[zdecode_line<T>(code) * g.nz + (code & 0xfff)] = {(T)v.x, (T)v.y};
  }
  __syncthreads();
  cpx<T>* r = fft_lds<T, -1>(a, b, pz, nl, g.nz, tw);
  const size_t plane = (size_t)g.nplane_x * g.ny;
  cpx<T>* zs = zbuf + (size_t)s * g.nz * plane;
  for (int idx = threadIdx.x; idx < nl * g.nz; idx += blockDim.x) {
    const int z = idx / nl, li = idx - z * nl;
    zs[(size_t)z * plane + line_dst[l0 + li]] = r[li * g.nz + z];
  }
}

// z FFT (FFTW_BACKWARD) + packGSpace + x(-2/N) (CP_State_GSpacePlane.C:1582-1627)
// + GStateSlab::addForces weights (stateSlab.C:240-272)
// + CP_eke_calc kinetic force and energy (cp_eke.C:72-130).
template <typename T>
__global__ __launch_bounds__(OA_NTHR) void k_zbwd(Geom g, FftPlan pz, int lpb,
                                                  const cpx<T>* __restrict__ zbuf,
                                                  const int* __restrict__ blk_pt0,
                                                  const int* __restrict__ pt_code,
                                                  const int* __restrict__ line_dst,
                                                  const double2* __restrict__ twz,
                                                  const double2* __restrict__ psi,
                                                  const double* __restrict__ g2,
                                                  const double* __restrict__ occ,
                                                  double ecut_ry, double scale,
                                                  double2* __restrict__ force,
                                                  double* __restrict__ eke_part) {
  extern __shared__ __align__(16) char smem_raw[];
  cpx<T>* tw = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* a = tw + g.nz;
  cpx<T>* b = a + lpb * g.nz;
  __shared__ double red[OA_NTHR / 64];
  const int s = blockIdx.y, blk = blockIdx.x;
  const int l0 = blk * lpb;
  const int nl = min(lpb, g.nlines - l0);
  load_tw(tw, twz, g.nz);
  const size_t plane = (size_t)g.nplane_x * g.ny;
  const cpx<T>* zs = zbuf + (size_t)s * g.nz * plane;
  for (int idx = threadIdx.x; idx < nl * g.nz; idx += blockDim.x) {
    const int z = idx / nl, li = idx - z * nl;
    a[li * g.nz + z] = zs[(size_t)z * plane + line_dst[l0 + li]];
  }
  __syncthreads();
  cpx<T>* r = fft_lds<T, +1>(a, b, pz, nl, g.nz, tw);
  const int p0 = blk_pt0[blk], p1 = blk_pt0[blk + 1];
  const double2* ps = psi + (size_t)s * g.ld;
  double2* fs = force + (size_t)s * g.ld;
  const double oc = occ[s];
  double eke = 0.0;
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const int code = pt_code[p];
    const cpx<T> v = r[zdecode_line<T>(code) * g.nz + (code & 0xfff)];
    const double w = (code < 0) ? 1.0 : 2.0;  // sign bit marks kx = 0
    const double gg = g2[p];
    double2 f;
    if (gg <= ecut_ry) {
      const double2 c = ps[p];
      const double wt = oc * w * gg;
      f.x = w * (scale * (double)v.x) - wt * c.x;
      f.y = w * (scale * (double)v.y) - wt * c.y;
      eke += wt * (c.x * c.x + c.y * c.y);
    } else {
      f.x = 0.0;
      f.y = 0.0;
    }
    fs[p] = f;
  }
  // workgroup reduction of eke (wave64 shuffles then LDS)
  for (int o = 32; o > 0; o >>= 1) eke += __shfl_down(eke, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = eke;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
    eke_part[(size_t)s * gridDim.x + blk] = 0.5 * t;
  }
}

// ---------------------------------------------------------------- y passes

// S = -1: y-FFT FFTW_FORWARD of the nplane_x populated kx columns
//         (fftCache.C:1180-1189), zbuf[kx][ky] -> xbuf[y][kx].
// S = +1: y-FFT FFTW_BACKWARD (fftCache.C:1294-1303), xbuf[y][kx] -> zbuf[kx][ky].
// Only ky inside the sphere's range for a column is read or written; the rest
// of the column is zero by construction.
template <typename T, int S>
__global__ __launch_bounds__(OA_NTHR) void k_y(Geom g, FftPlan py, int kc,
                                               const int* __restrict__ kylo,
                                               const int* __restrict__ kyhi,
                                               const double2* __restrict__ twy,
                                               const cpx<T>* __restrict__ src,
                                               cpx<T>* __restrict__ dst) {
  extern __shared__ __align__(16) char smem_raw[];
  cpx<T>* tw = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* a = tw + g.ny;
  cpx<T>* b = a + kc * g.ny;
  const int kx0 = blockIdx.x * kc;
  const int nk = min(kc, g.nplane_x - kx0);
  const int z = blockIdx.y, s = blockIdx.z;
  load_tw(tw, twy, g.ny);
  const size_t plane = (size_t)g.nplane_x * g.ny;
  const size_t off = ((size_t)s * g.nz + z) * plane;
  const int ny = g.ny;
  if (S < 0) {
    for (int idx = threadIdx.x; idx < nk * ny; idx += blockDim.x) {
      const int kxl = idx / ny, y = idx - kxl * ny;
      const int kx = kx0 + kxl;
      const int kyv = (y <= (ny - 1) / 2) ? y : y - ny;
      a[kxl * ny + y] = (kyv >= kylo[kx] && kyv <= kyhi[kx]) ? src[off + (size_t)kx * ny + y]
                                                              : cpx<T>{(T)0, (T)0};
    }
  } else {
    for (int idx = threadIdx.x; idx < nk * ny; idx += blockDim.x) {
      const int y = idx / nk, kxl = idx - y * nk;
      a[kxl * ny + y] = src[off + (size_t)y * g.nplane_x + kx0 + kxl];
    }
  }
  __syncthreads();
  cpx<T>* r = fft_lds<T, S>(a, b, py, nk, ny, tw);
  if (S < 0) {
    for (int idx = threadIdx.x; idx < nk * ny; idx += blockDim.x) {
      const int y = idx / nk, kxl = idx - y * nk;
      dst[off + (size_t)y * g.nplane_x + kx0 + kxl] = r[kxl * ny + y];
    }
  } else {
    for (int idx = threadIdx.x; idx < nk * ny; idx += blockDim.x) {
      const int kxl = idx / ny, y = idx - kxl * ny;
      const int kx = kx0 + kxl;
      const int kyv = (y <= (ny - 1) / 2) ? y : y - ny;
      if (kyv >= kylo[kx] && kyv <= kyhi[kx]) dst[off + (size_t)kx * ny + y] = r[kxl * ny + y];
    }
  }
}

// ---------------------------------------------------------------- x passes

// conjugate + FFTW c2r of each row (fftCache.C:1207-1241) done as one forward
// complex FFT of two hermitian-extended rows (Y1 + i Y2); then the density
// contribution psi^2 occ / Omega summed over states in registers
// (CP_State_RealSpacePlane.C:349-355, CP_Rho_RealSpacePlane.C:630-650).
template <typename T>
__global__ __launch_bounds__(OA_NTHR) void k_xc2r(Geom g, FftPlan px, int npair,
                                                  const double2* __restrict__ twx,
                                                  const cpx<T>* __restrict__ xbuf,
                                                  const double* __restrict__ occ, double inv_vol,
                                                  T* __restrict__ psir,
                                                  double* __restrict__ rho) {
  extern __shared__ __align__(16) char smem_raw[];
  cpx<T>* tw = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* a = tw + g.nx;
  cpx<T>* b = a + npair * g.nx;
  const int nx = g.nx, ny = g.ny, npx = g.nplane_x;
  const int y0 = blockIdx.x * 2 * npair, z = blockIdx.y;
  load_tw(tw, twx, nx);
  double acc0[OA_XSLOTS], acc1[OA_XSLOTS];
#pragma unroll
  for (int i = 0; i < OA_XSLOTS; i++) { acc0[i] = 0.0; acc1[i] = 0.0; }
  const int nslot = npair * nx;
  const size_t N = (size_t)nx * ny * g.nz;
  for (int s = 0; s < g.nst; s++) {
    const cpx<T>* xs = xbuf + ((size_t)s * g.nz + z) * ny * npx;
    for (int idx = threadIdx.x; idx < nslot; idx += blockDim.x) {
      const int q = idx / nx, k = idx - q * nx;
      const int y1 = y0 + 2 * q, y2 = y1 + 1;
      cpx<T> Z = {(T)0, (T)0};
      int kk = -1;
      bool mirror = false;
      if (k < npx) kk = k;
      else if (k > nx - npx) { kk = nx - k; mirror = true; }
      if (kk >= 0 && y1 < ny) {
        cpx<T> X1 = xs[(size_t)y1 * npx + kk];
        cpx<T> X2 = (y2 < ny) ? xs[(size_t)y2 * npx + kk] : cpx<T>{(T)0, (T)0};
        if (kk == 0) Z = {X1.x, X2.x};
        else if (!mirror) Z = {X1.x - X2.y, X1.y + X2.x};
        else Z = {X1.x + X2.y, X2.x - X1.y};
      }
      a[idx] = Z;
    }
    __syncthreads();
    cpx<T>* r = fft_lds<T, -1>(a, b, px, npair, nx, tw);
    const double oc = occ[s];
    T* pr = psir + (size_t)s * N + (size_t)z * ny * nx;
#pragma unroll
    for (int i = 0; i < OA_XSLOTS; i++) {
      const int idx = threadIdx.x + i * OA_NTHR;
      if (idx < nslot) {
        const int q = idx / nx, x = idx - q * nx;
        const int y1 = y0 + 2 * q, y2 = y1 + 1;
        const cpx<T> v = r[idx];
        if (y1 < ny) {
          pr[(size_t)y1 * nx + x] = v.x;
          acc0[i] += oc * (double)v.x * (double)v.x;
        }
        if (y2 < ny) {
          pr[(size_t)y2 * nx + x] = v.y;
          acc1[i] += oc * (double)v.y * (double)v.y;
        }
      }
    }
    __syncthreads();
  }
  double* rz = rho + (size_t)z * ny * nx;
#pragma unroll
  for (int i = 0; i < OA_XSLOTS; i++) {
    const int idx = threadIdx.x + i * OA_NTHR;
    if (idx < nslot) {
      const int q = idx / nx, x = idx - q * nx;
      const int y1 = y0 + 2 * q, y2 = y1 + 1;
      if (y1 < ny) rz[(size_t)y1 * nx + x] = acc0[i] * inv_vol;
      if (y2 < ny) rz[(size_t)y2 * nx + x] = acc1[i] * inv_vol;
    }
  }
}

// V(r) psi(r) occ product (CP_State_RealSpacePlane.C:580-596), then FFTW r2c and
// the conjugation of fftCache.C:1263-1274, two real rows per complex FFT.
template <typename T>
__global__ __launch_bounds__(OA_NTHR) void k_xr2c(Geom g, FftPlan px, int npair,
                                                  const double2* __restrict__ twx,
                                                  const T* __restrict__ psir,
                                                  const double* __restrict__ vks,
                                                  const double* __restrict__ occ,
                                                  cpx<T>* __restrict__ xbuf) {
  extern __shared__ __align__(16) char smem_raw[];
  cpx<T>* tw = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* a = tw + g.nx;
  cpx<T>* b = a + npair * g.nx;
  const int nx = g.nx, ny = g.ny, npx = g.nplane_x;
  const int y0 = blockIdx.x * 2 * npair, z = blockIdx.y;
  load_tw(tw, twx, nx);
  const int nslot = npair * nx;
  const size_t N = (size_t)nx * ny * g.nz;
  T v1[OA_XSLOTS], v2[OA_XSLOTS];
  const double* vz = vks + (size_t)z * ny * nx;
#pragma unroll
  for (int i = 0; i < OA_XSLOTS; i++) {
    const int idx = threadIdx.x + i * OA_NTHR;
    v1[i] = (T)0;
    v2[i] = (T)0;
    if (idx < nslot) {
      const int q = idx / nx, x = idx - q * nx;
      const int y1 = y0 + 2 * q, y2 = y1 + 1;
      if (y1 < ny) v1[i] = (T)vz[(size_t)y1 * nx + x];
      if (y2 < ny) v2[i] = (T)vz[(size_t)y2 * nx + x];
    }
  }
  const int nout = npair * npx;
  for (int s = 0; s < g.nst; s++) {
    const T oc = (T)occ[s];
    const T* pr = psir + (size_t)s * N + (size_t)z * ny * nx;
#pragma unroll
    for (int i = 0; i < OA_XSLOTS; i++) {
      const int idx = threadIdx.x + i * OA_NTHR;
      if (idx < nslot) {
        const int q = idx / nx, x = idx - q * nx;
        const int y1 = y0 + 2 * q, y2 = y1 + 1;
        T f1 = (y1 < ny) ? pr[(size_t)y1 * nx + x] * (v1[i] * oc) : (T)0;
        T f2 = (y2 < ny) ? pr[(size_t)y2 * nx + x] * (v2[i] * oc) : (T)0;
        a[idx] = {f1, f2};
      }
    }
    __syncthreads();
    cpx<T>* r = fft_lds<T, -1>(a, b, px, npair, nx, tw);
    cpx<T>* xs = xbuf + ((size_t)s * g.nz + z) * ny * npx;
    for (int idx = threadIdx.x; idx < nout; idx += blockDim.x) {
      const int q = idx / npx, k = idx - q * npx;
      const int y1 = y0 + 2 * q, y2 = y1 + 1;
      const cpx<T> Zk = r[q * nx + k];
      const cpx<T> Zm = r[q * nx + (k == 0 ? 0 : nx - k)];
      // X1 = (Zk + conj Zm)/2, X2 = (Zk - conj Zm)/(2i); store conj(X)
      const cpx<T> c1 = {(T)0.5 * (Zk.x + Zm.x), -(T)0.5 * (Zk.y - Zm.y)};
      const cpx<T> c2 = {(T)0.5 * (Zk.y + Zm.y), (T)0.5 * (Zk.x - Zm.x)};
      if (y1 < ny) xs[(size_t)y1 * npx + k] = c1;
      if (y2 < ny) xs[(size_t)y2 * npx + k] = c2;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- fused planes
//
// One workgroup owns one z-plane for a chunk of `cs` states and keeps the
// whole half plane Y[kx][y] (nplane_x x Ny complex) in LDS, so the y-FFT and
// the x-FFT of a state meet in LDS instead of in HBM.  Every thread owns the
// fixed plane positions r = tid + i*NTH (i < MAXS) for the whole launch:
// that is where rho is accumulated (forward) and V is held (backward).  A
// batch of pair-rows [ya, yb) covers the positions [ya*nx, yb*nx), so a slot
// is live in a batch iff its position falls in that range.

// y-FFT FFTW_FORWARD on the populated columns, conjugate + c2r per row as one
// forward FFT of two Hermitian-extended rows, psi(r) out, sum occ psi^2 in
// registers (fftCache.C:1167-1243; CP_State_RealSpacePlane.C:349-355).
template <typename T, int NTH, int MAXS>
__global__ __launch_bounds__(NTH) void k_plane_fwd(Geom g, FftPlan px, FftPlan py, int kc, int pc,
                                                   int cs, const int* __restrict__ kylo,
                                                   const int* __restrict__ kyhi,
                                                   const double2* __restrict__ twx,
                                                   const double2* __restrict__ twy,
                                                   const cpx<T>* __restrict__ zbuf,
                                                   const double* __restrict__ occ,
                                                   T* __restrict__ psir,
                                                   double* __restrict__ rho_part) {
  extern __shared__ __align__(16) char smem_raw[];
  const int nx = g.nx, ny = g.ny, npx = g.nplane_x;
  const int scr = max(kc * ny, pc * nx);
  cpx<T>* twX = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* twY = twX + nx;
  cpx<T>* Y = twY + ny;
  cpx<T>* s1 = Y + npx * ny;
  cpx<T>* s2 = s1 + scr;
  const int z = blockIdx.x, chunk = blockIdx.y;
  const int st0 = chunk * cs, st1 = min(g.nst, st0 + cs);
  const int NP = nx * ny;
  const size_t N = (size_t)NP * g.nz;
  const size_t plane = (size_t)npx * ny;
  load_tw(twX, twx, nx);
  load_tw(twY, twy, ny);
  double acc[MAXS];
#pragma unroll
  for (int i = 0; i < MAXS; i++) acc[i] = 0.0;
  const int npairs = (ny + 1) / 2;
  for (int s = st0; s < st1; s++) {
    // ---- y pass: load the populated columns, FFT in column chunks
    const cpx<T>* zs = zbuf + ((size_t)s * g.nz + z) * plane;
    for (int idx = threadIdx.x; idx < npx * ny; idx += NTH) {
      const int kx = idx / ny, y = idx - kx * ny;
      const int kyv = (y <= (ny - 1) / 2) ? y : y - ny;
      Y[idx] = (kyv >= kylo[kx] && kyv <= kyhi[kx]) ? zs[idx] : cpx<T>{(T)0, (T)0};
    }
    __syncthreads();
    for (int kx0 = 0; kx0 < npx; kx0 += kc) {
      const int nk = min(kc, npx - kx0);
      cpx<T>* a = Y + kx0 * ny;
      cpx<T>* r = fft_lds<T, -1>(a, s1, py, nk, ny, twY);
      if (r != a) {
        for (int i = threadIdx.x; i < nk * ny; i += NTH) a[i] = r[i];
        __syncthreads();
      }
    }
    // ---- x pass in chunks of pair-rows
    const double oc = occ[s];
    T* pr = psir + (size_t)s * N + (size_t)z * NP;
    for (int q0 = 0; q0 < npairs; q0 += pc) {
      const int np = min(pc, npairs - q0);
      for (int idx = threadIdx.x; idx < np * nx; idx += NTH) {
        const int ql = idx / nx, k = idx - ql * nx;
        const int y1 = 2 * (q0 + ql), y2 = y1 + 1;
        cpx<T> Z = {(T)0, (T)0};
        int kk = -1;
        bool mirror = false;
        if (k < npx) kk = k;
        else if (k > nx - npx) { kk = nx - k; mirror = true; }
        if (kk >= 0) {
          const cpx<T> X1 = Y[kk * ny + y1];
          const cpx<T> X2 = (y2 < ny) ? Y[kk * ny + y2] : cpx<T>{(T)0, (T)0};
          if (kk == 0) Z = {X1.x, X2.x};
          else if (!mirror) Z = {X1.x - X2.y, X1.y + X2.x};
          else Z = {X1.x + X2.y, X2.x - X1.y};
        }
        s1[idx] = Z;
      }
      __syncthreads();
      cpx<T>* r = fft_lds<T, -1>(s1, s2, px, np, nx, twX);
      const int ya = 2 * q0, yb = min(ny, 2 * (q0 + np));
      const int ra = ya * nx, rb = yb * nx;
#pragma unroll
      for (int i = 0; i < MAXS; i++) {
        const int rr = threadIdx.x + i * NTH;
        if (rr >= ra && rr < rb) {
          const int off = rr - ra;
          const int yl = off / nx;  // ya is even: yl carries the row parity
          const int x = off - yl * nx;
          const cpx<T> v = r[(yl >> 1) * nx + x];
          const T val = (yl & 1) ? v.y : v.x;
          pr[rr] = val;
          acc[i] += oc * (double)val * (double)val;
        }
      }
      __syncthreads();
    }
  }
  double* rp = rho_part + (size_t)chunk * N + (size_t)z * NP;
#pragma unroll
  for (int i = 0; i < MAXS; i++) {
    const int rr = threadIdx.x + i * NTH;
    if (rr < NP) rp[rr] = acc[i];
  }
}

// V psi occ product, FFTW r2c + conjugate per row (two real rows per complex
// FFT), y-FFT FFTW_BACKWARD on the populated columns
// (CP_State_RealSpacePlane.C:580-596; fftCache.C:1252-1319).
template <typename T, int NTH, int MAXS>
__global__ __launch_bounds__(NTH) void k_plane_bwd(Geom g, FftPlan px, FftPlan py, int kc, int pc,
                                                   int cs, const int* __restrict__ kylo,
                                                   const int* __restrict__ kyhi,
                                                   const double2* __restrict__ twx,
                                                   const double2* __restrict__ twy,
                                                   const T* __restrict__ psir,
                                                   const double* __restrict__ vks,
                                                   const double* __restrict__ occ,
                                                   cpx<T>* __restrict__ zbuf) {
  extern __shared__ __align__(16) char smem_raw[];
  const int nx = g.nx, ny = g.ny, npx = g.nplane_x;
  const int scr = max(kc * ny, pc * nx);
  cpx<T>* twX = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* twY = twX + nx;
  cpx<T>* Y = twY + ny;
  cpx<T>* s1 = Y + npx * ny;
  cpx<T>* s2 = s1 + scr;
  const int z = blockIdx.x, chunk = blockIdx.y;
  const int st0 = chunk * cs, st1 = min(g.nst, st0 + cs);
  const int NP = nx * ny;
  const size_t N = (size_t)NP * g.nz;
  const size_t plane = (size_t)npx * ny;
  load_tw(twX, twx, nx);
  load_tw(twY, twy, ny);
  T vv[MAXS];
  const double* vz = vks + (size_t)z * NP;
#pragma unroll
  for (int i = 0; i < MAXS; i++) {
    const int rr = threadIdx.x + i * NTH;
    vv[i] = (rr < NP) ? (T)vz[rr] : (T)0;
  }
  const int npairs = (ny + 1) / 2;
  for (int s = st0; s < st1; s++) {
    const T oc = (T)occ[s];
    const T* pr = psir + (size_t)s * N + (size_t)z * NP;
    // ---- x pass: rows -> populated kx of the half plane Y[kx][y] (conjugated)
    for (int q0 = 0; q0 < npairs; q0 += pc) {
      const int np = min(pc, npairs - q0);
      const int ya = 2 * q0, yb = min(ny, 2 * (q0 + np));
      const int ra = ya * nx, rb = yb * nx;
#pragma unroll
      for (int i = 0; i < MAXS; i++) {
        const int rr = threadIdx.x + i * NTH;
        if (rr >= ra && rr < rb) {
          const int off = rr - ra;
          const int yl = off / nx;
          const int x = off - yl * nx;
          const T f = pr[rr] * (vv[i] * oc);
          T* dst = reinterpret_cast<T*>(&s1[(yl >> 1) * nx + x]);
          dst[yl & 1] = f;
        }
      }
      if ((ny & 1) && q0 + np == npairs) {
        // the last pair of an odd Ny has no second row
        for (int x = threadIdx.x; x < nx; x += NTH) s1[(np - 1) * nx + x].y = (T)0;
      }
      __syncthreads();
      cpx<T>* r = fft_lds<T, -1>(s1, s2, px, np, nx, twX);
      for (int idx = threadIdx.x; idx < np * npx; idx += NTH) {
        const int ql = idx / npx, k = idx - ql * npx;
        const int y1 = ya + 2 * ql, y2 = y1 + 1;
        const cpx<T> Zk = r[ql * nx + k];
        const cpx<T> Zm = r[ql * nx + (k == 0 ? 0 : nx - k)];
        const cpx<T> c1 = {(T)0.5 * (Zk.x + Zm.x), -(T)0.5 * (Zk.y - Zm.y)};
        const cpx<T> c2 = {(T)0.5 * (Zk.y + Zm.y), (T)0.5 * (Zk.x - Zm.x)};
        Y[k * ny + y1] = c1;
        if (y2 < ny) Y[k * ny + y2] = c2;
      }
      __syncthreads();
    }
    // ---- y pass backward, column chunks, populated ky written out
    cpx<T>* zs = zbuf + ((size_t)s * g.nz + z) * plane;
    for (int kx0 = 0; kx0 < npx; kx0 += kc) {
      const int nk = min(kc, npx - kx0);
      cpx<T>* a = Y + kx0 * ny;
      cpx<T>* r = fft_lds<T, +1>(a, s1, py, nk, ny, twY);
      for (int idx = threadIdx.x; idx < nk * ny; idx += NTH) {
        const int kxl = idx / ny, y = idx - kxl * ny;
        const int kx = kx0 + kxl;
        const int kyv = (y <= (ny - 1) / 2) ? y : y - ny;
        if (kyv >= kylo[kx] && kyv <= kyhi[kx]) zs[(size_t)kx * ny + y] = r[idx];
      }
      __syncthreads();
    }
  }
}

// rho = sum of the chunk partials / Omega (added to rho when accumulate is
// set), and the partials of sum V rho when vks is set.  Two points per thread and round, all their loads issued
// before any is used (the partials are read once: latency, not bandwidth,
// bounds a plain loop).
__global__ void k_rho_reduce(const double* __restrict__ part, int nchunk, size_t n,
                             double inv_vol, const double* __restrict__ vks,
                             double* __restrict__ rho, double* __restrict__ dpart,
                             int accumulate) {
  __shared__ double red[16];
  double e = 0.0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t r0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x; r0 < n; r0 += 2 * stride) {
    const size_t r1 = r0 + stride, r1c = r1 < n ? r1 : r0;
    double a0 = 0.0, a1 = 0.0;
    int c = 0;
    for (; c + 4 <= nchunk; c += 4) {
      double v0[4], v1[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        v0[u] = part[(size_t)(c + u) * n + r0];
        v1[u] = part[(size_t)(c + u) * n + r1c];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        a0 += v0[u];
        a1 += v1[u];
      }
    }
    for (; c < nchunk; c++) {
      a0 += part[(size_t)c * n + r0];
      a1 += part[(size_t)c * n + r1c];
    }
    a0 *= inv_vol;
    a1 *= inv_vol;
    if (accumulate) {
      a0 += rho[r0];
      a1 += rho[r1c];
    }
    rho[r0] = a0;
    if (r1 < n) rho[r1] = a1;
    if (vks) {
      e += a0 * vks[r0];
      if (r1 < n) e += a1 * vks[r1];
    }
  }
  for (int o = 32; o > 0; o >>= 1) e += __shfl_down(e, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
    dpart[blockIdx.x] = t;
  }
}

// ---------------------------------------------------------------- launchers

static size_t fft_smem(int n, int batch, size_t csz) { return csz * ((size_t)n + 2 * (size_t)batch * n); }

size_t plane_smem(const FftCtx& c, bool fp32) {
  const size_t cs = fp32 ? 8 : 16;
  const Geom& g = c.g;
  const size_t scr = (size_t)std::max(c.kcp * g.ny, c.pcp * g.nx);
  return cs * ((size_t)g.nx + g.ny + (size_t)g.nplane_x * g.ny + 2 * scr);
}

template <typename T, int NTH, int MAXS>
hipError_t launch_plane(const FftCtx& c, bool fwd, hipStream_t st) {
  const Geom& g = c.g;
  const dim3 grid(g.nz, c.nchunk);
  const size_t smem = plane_smem(c, sizeof(T) == 4);
  if (fwd) {
    auto k = k_plane_fwd<T, NTH, MAXS>;
    if (hipError_t e = grant_lds((const void*)k, smem); e != hipSuccess) return e;
    hipLaunchKernelGGL(k, grid, dim3(NTH), smem, st, g, c.px, c.py, c.kcp, c.pcp, c.cs, c.kylo,
                       c.kyhi, c.twx, c.twy, (const cpx<T>*)c.zbuf, c.occ, (T*)c.psir,
                       c.rho_part);
  } else {
    auto k = k_plane_bwd<T, NTH, MAXS>;
    if (hipError_t e = grant_lds((const void*)k, smem); e != hipSuccess) return e;
    hipLaunchKernelGGL(k, grid, dim3(NTH), smem, st, g, c.px, c.py, c.kcp, c.pcp, c.cs, c.kylo,
                       c.kyhi, c.twx, c.twy, (const T*)c.psir, c.vks, c.occ, (cpx<T>*)c.zbuf);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_plane_sel(const FftCtx& c, bool fwd, hipStream_t st) {
  switch (c.plane_cfg) {
    case 0: return launch_plane<T, 1024, 4>(c, fwd, st);
    case 1: return launch_plane<T, 1024, 10>(c, fwd, st);
    case 2: return launch_plane<T, 1024, 16>(c, fwd, st);
    default: return hipErrorInvalidValue;
  }
}

// register-slot variant for a plane of np points (1024 threads each)
int plane_config_for(int np) {
  if (np <= 1024 * 4) return 0;
  if (np <= 1024 * 10) return 1;
  if (np <= 1024 * 16) return 2;
  return -1;
}

int plane_threads(int) { return 1024; }

hipError_t plane_fwd(const FftCtx& c, bool fp32, hipStream_t st) {
  return fp32 ? launch_plane_sel<float>(c, true, st) : launch_plane_sel<double>(c, true, st);
}

hipError_t plane_bwd(const FftCtx& c, bool fp32, hipStream_t st) {
  return fp32 ? launch_plane_sel<float>(c, false, st) : launch_plane_sel<double>(c, false, st);
}

hipError_t rho_reduce(const FftCtx& c, double* dpart, int ndpart, hipStream_t st, bool with_v,
                      bool accumulate) {
  const size_t n = (size_t)c.g.nx * c.g.ny * c.g.nz;
  hipLaunchKernelGGL(k_rho_reduce, dim3(ndpart), dim3(256), 0, st, c.rho_part, c.nchunk, n,
                     c.inv_vol, with_v ? c.vks : nullptr, c.rho, dpart, accumulate ? 1 : 0);
  return hipGetLastError();
}

// Stage s of the chain: 0 k_zfwd, 1 k_y<-1>, 2 k_xc2r, 3 k_xr2c, 4 k_y<+1>, 5 k_zbwd.
template <typename T>
hipError_t launch_stage(const FftCtx& c, int stage, hipStream_t st) {
  const Geom& g = c.g;
  const size_t cs = sizeof(cpx<T>);
  const dim3 zgrid((g.nlines + c.lpb - 1) / c.lpb, g.nst);
  const dim3 ygrid((g.nplane_x + c.kc - 1) / c.kc, g.nz, g.nst);
  const dim3 xgrid((g.ny + 2 * c.npair - 1) / (2 * c.npair), g.nz);
  switch (stage) {
    case 0:
      hipLaunchKernelGGL(k_zfwd<T>, zgrid, dim3(OA_NTHR), fft_smem(g.nz, c.lpb, cs), st, g, c.pz,
                         c.lpb, c.psi, c.blk_pt0, c.pt_code, c.line_dst, c.twz,
                         (cpx<T>*)c.zbuf);
      break;
    case 1:
      hipLaunchKernelGGL((k_y<T, -1>), ygrid, dim3(OA_NTHR), fft_smem(g.ny, c.kc, cs), st, g,
                         c.py, c.kc, c.kylo, c.kyhi, c.twy, (const cpx<T>*)c.zbuf,
                         (cpx<T>*)c.xbuf);
      break;
    case 2:
      hipLaunchKernelGGL(k_xc2r<T>, xgrid, dim3(OA_NTHR), fft_smem(g.nx, c.npair, cs), st, g,
                         c.px, c.npair, c.twx, (const cpx<T>*)c.xbuf, c.occ, c.inv_vol,
                         (T*)c.psir, c.rho);
      break;
    case 3:
      hipLaunchKernelGGL(k_xr2c<T>, xgrid, dim3(OA_NTHR), fft_smem(g.nx, c.npair, cs), st, g,
                         c.px, c.npair, c.twx, (const T*)c.psir, c.vks, c.occ,
                         (cpx<T>*)c.xbuf);
      break;
    case 4:
      hipLaunchKernelGGL((k_y<T, +1>), ygrid, dim3(OA_NTHR), fft_smem(g.ny, c.kc, cs), st, g,
                         c.py, c.kc, c.kylo, c.kyhi, c.twy, (const cpx<T>*)c.xbuf,
                         (cpx<T>*)c.zbuf);
      break;
    case 5:
      hipLaunchKernelGGL(k_zbwd<T>, zgrid, dim3(OA_NTHR), fft_smem(g.nz, c.lpb, cs), st, g, c.pz,
                         c.lpb, (const cpx<T>*)c.zbuf, c.blk_pt0, c.pt_code, c.line_dst, c.twz,
                         c.psi, c.g2, c.occ, c.ecut_ry, c.scale, c.force, c.eke_part);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t fft_stage(const FftCtx& c, bool fp32, int stage, hipStream_t st) {
  return fp32 ? launch_stage<float>(c, stage, st) : launch_stage<double>(c, stage, st);
}

size_t fft_max_smem(const FftCtx& c, bool fp32) {
  const size_t cs = fp32 ? 8 : 16;
  size_t m = fft_smem(c.g.nz, c.lpb, cs);
  size_t t = fft_smem(c.g.ny, c.kc, cs);
  if (t > m) m = t;
  t = fft_smem(c.g.nx, c.npair, cs);
  if (t > m) m = t;
  return m;
}

}  // namespace oa
