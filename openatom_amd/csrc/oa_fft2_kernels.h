// oa_fft2_kernels.h -- the six passes of the state FFT chain on the
// register-resident two-factor engine (oa_fft2.h).  Each kernel is a
// template on the transform length N = R1*R2 of its axis; oa_fft2_inst.hip
// instantiates one length per translation unit.
//
//   k2_zfwd : sphere gather + z-FFT (FFTW_FORWARD)  psi(G) -> zbuf[s][z][kx][ky]
//   k2_yfwd : y-FFT (FFTW_FORWARD), transposing     zbuf   -> xbuf[s][z][y][kx]
//   k2_xc2r : pair-row Hermitian x-FFT, rho         xbuf   -> psir, rho partials
//   k2_xr2c : V psi occ, pair-row x-FFT             psir,V -> xbuf (conjugated)
//   k2_ybwd : y-FFT (FFTW_BACKWARD)                 xbuf   -> zbuf
//   k2_zbwd : z-FFT, pack, -2/N, kinetic force, EKE zbuf   -> F(G)
// Conventions are those of fftCache.C:460-553, 1099-1319 and
// CP_State_GSpacePlane.C:1582-1778 (see oa_fft.hip for the restatement).
#pragma once

#include "oa_fft2.h"
#include "oa_kernels.h"

namespace oa2 {

using oa::FftCtx;

constexpr int kThreads = 256;

template <int R1, int R2>
struct Shape {
  static constexpr int N = R1 * R2;
  static constexpr int G = R1 > R2 ? R1 : R2;  // threads per transform
  static constexpr int NTR = kThreads / G;     // transforms per workgroup
};

__device__ __forceinline__ int signed_k(int i, int n) { return (i <= (n - 1) / 2) ? i : i - n; }

// XCD-aware block order of a 2D grid.  The chip hands workgroups to its 8
// XCDs (each with its own L2) round-robin by linear block id, so blocks x
// and x+1 would run on different XCDs.  This map gives each XCD a
// contiguous range of logical blocks (x fastest): neighbouring line blocks,
// which share the 128-byte lines at their edges, then meet in one L2.
constexpr int kXcds = 8;
__device__ __forceinline__ void xcd_block(int& bx, int& by) {
  const int nx = gridDim.x, total = nx * gridDim.y;
  const int lin = blockIdx.x + blockIdx.y * nx;
  const int xcd = lin % kXcds, k = lin / kXcds;
  const int q = total / kXcds, r = total % kXcds;
  const int logical = xcd < r ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
  bx = logical % nx;
  by = logical / nx;
}

// ------------------------------------------------------------------ z passes

// PR: every stored |kz| <= (N - 1) / 4 (pruned pass 1, Sph in oa_fft2.h)
template <typename T, int R1, int R2, bool PR>
__global__ __launch_bounds__(kThreads) void k2_zfwd(FftCtx c) {
  using Sh = Shape<R1, R2>;
  constexpr int N = Sh::N, G = Sh::G, NTR = Sh::NTR;
  __shared__ cpx<T> ex[NTR * N];
  int bx, s;
  xcd_block(bx, s);
  const int l0 = bx * NTR;
  const int nl = min(NTR, c.g.nlines - l0);
  const int tid = threadIdx.x;
  // pass 2's zbuf column, loaded up front (behind pass 1)
  const int tr2 = tid % NTR, j2 = tid / NTR;
  const int w2 = c.lines[l0 + min(tr2, nl - 1)].w;
  {  // pass 1: lanes along z (contiguous sphere points)
    const int tr = tid / G, j = tid - tr * G;
    if (tr < nl && j < R2) {
      const int4 L = c.lines[l0 + tr];
      const double2* ps = c.psi + (size_t)s * c.g.ld + L.x;
      // branch-free: every load issued before any is waited for (a load in
      // a divergent branch is waited for on its own); lines hold >= 1 point
      cpx<T> v[R1];
#pragma unroll
      for (int t = 0; t < R1; t++) {
        if (Sph<N, R2>::cut(PR, t)) {
          v[t] = {(T)0, (T)0};
          continue;
        }
        const int d = signed_k(j + t * R2, N) - L.z;
        const double2 a = ps[min(max(d, 0), L.y - 1)];
        const bool in = d >= 0 && d < L.y;
        v[t] = {in ? (T)a.x : (T)0, in ? (T)a.y : (T)0};
      }
      Dft<T, R1, -1>::run(v);
#pragma unroll
      for (int t = 0; t < R1; t++) ex[tr * N + j * R1 + t] = v[t];
    }
  }
  __syncthreads();
  {  // pass 2: lanes along lines (adjacent (kx,ky) columns of zbuf)
    const int tr = tr2, j = j2;
    if (tr < nl && j < R1) {
      cpx<T> u[R2];
#pragma unroll
      for (int t = 0; t < R2; t++) u[t] = ex[tr * N + j + t * R1];
#pragma unroll
      for (int t = 1; t < R2; t++) u[t] = cm(u[t], twiddle<T, -1>(c.twz, t * j));
      Dft<T, R2, -1>::run(u);
      const size_t plane = (size_t)c.g.nplane_x * c.g.ny;
      cpx<T>* zs = (cpx<T>*)c.zbuf + (size_t)s * c.g.nz * plane + w2;
#pragma unroll
      for (int t = 0; t < R2; t++) zs[(size_t)(j + t * R1) * plane] = u[t];
    }
  }
}

// PR: every stored |kz| <= (N - 1) / 4 (pass-2 outputs outside it dropped)
template <typename T, int R1, int R2, bool PR>
__global__ __launch_bounds__(kThreads) void k2_zbwd(FftCtx c) {
  using Sh = Shape<R1, R2>;
  constexpr int N = Sh::N, G = Sh::G, NTR = Sh::NTR;
  constexpr int kPre = 4;  // passes over the points whose operands are loaded up front
  __shared__ cpx<T> ex[NTR * N];
  __shared__ unsigned char wt_l[NTR * N];  // per point: 1 on kx = 0 lines, 2 elsewhere
  __shared__ double red[kThreads / 64];
  int bx, s;
  xcd_block(bx, s);
  const int l0 = bx * NTR;
  const int nl = min(NTR, c.g.nlines - l0);
  const int tid = threadIdx.x;
  const size_t plane = (size_t)c.g.nplane_x * c.g.ny;
  // The workgroup's lines are one contiguous run of stored points [p0, p1):
  // the pack stage reads psi and |G|^2 and writes F along it, lanes on
  // consecutive points.  Lanes along the z of each line would store runs
  // of R1 points at arbitrary alignment, partial sectors of HBM that cost
  // more traffic than the data.
  const int p0 = c.lines[l0].x;
  const int4 Llast = c.lines[l0 + nl - 1];
  const int np = Llast.x + Llast.y - p0;  // <= nl * N
  const double2* ps = c.psi + (size_t)s * c.g.ld + p0;
  double2* fs = c.force + (size_t)s * c.g.ld + p0;
  const double* g2 = c.g2 + p0;
  // pass 2's twiddles and line, and the first pack operands, loaded up
  // front (branch-free, clamped) so that their latency hides behind pass 1
  const int tr2 = tid / G, j2 = tid - tr2 * G;
  cpx<T> tw2[R2];
#pragma unroll
  for (int t = 1; t < R2; t++) tw2[t] = twiddle<T, +1>(c.twz, t * min(j2, R1 - 1));
  const int4 L = c.lines[l0 + min(tr2, nl - 1)];
  const int w1 = c.lines[l0 + min(tid % NTR, nl - 1)].w;
  const double oc = c.occ[s];
  double gq[kPre];
  double2 cq[kPre];
#pragma unroll
  for (int i = 0; i < kPre; i++) {
    const int o = min(tid + i * kThreads, np - 1);
    gq[i] = g2[o];
    cq[i] = ps[o];
  }
  {  // pass 1: lanes along lines
    const int tr = tid % NTR, j = tid / NTR;
    if (tr < nl && j < R2) {
      const cpx<T>* zs = (const cpx<T>*)c.zbuf + (size_t)s * c.g.nz * plane + w1;
      cpx<T> v[R1];
#pragma unroll
      for (int t = 0; t < R1; t++) v[t] = zs[(size_t)(j + t * R2) * plane];
      Dft<T, R1, +1>::run(v);
#pragma unroll
      for (int t = 0; t < R1; t++) ex[tr * N + j * R1 + t] = v[t];
    }
  }
  __syncthreads();
  {  // pass 2: lanes along z; the results go back to LDS in stored-point order
    const int tr = tr2, j = j2;
    const bool act = tr < nl && j < R1;
    cpx<T> u[R2];
    if (act) {
#pragma unroll
      for (int t = 0; t < R2; t++) u[t] = ex[tr * N + j + t * R1];
#pragma unroll
      for (int t = 1; t < R2; t++) u[t] = cm(u[t], tw2[t]);
      Dft<T, R2, +1>::run(u);
    }
    __syncthreads();  // ex is rewritten below
    if (act) {
      const int ob = L.x - p0;
      const unsigned char w = (L.w < c.g.ny) ? 1 : 2;  // kx = 0 lines have dst < ny
#pragma unroll
      for (int t = 0; t < R2; t++) {
        const int d = signed_k(j + t * R1, N) - L.z;
        if (!Sph<N, R1>::cut(PR, t) && d >= 0 && d < L.y) {
          ex[ob + d] = u[t];
          wt_l[ob + d] = w;
        }
      }
    }
  }
  __syncthreads();
  // pack: -2/N scale (w doubles the kx > 0 half), kinetic force, EKE
  double eke = 0.0;
  auto pack = [&](int o, double gg, double2 cc) __attribute__((always_inline)) {
    const cpx<T> u = ex[o];
    const double w = (double)wt_l[o];
    double2 f;
    if (gg <= c.ecut_ry) {
      const double wt = oc * w * gg;
      f.x = w * (c.scale * (double)u.x) - wt * cc.x;
      f.y = w * (c.scale * (double)u.y) - wt * cc.y;
      eke += wt * (cc.x * cc.x + cc.y * cc.y);
    } else {
      f.x = 0.0;
      f.y = 0.0;
    }
    fs[o] = f;
  };
#pragma unroll
  for (int i = 0; i < kPre; i++) {
    const int o = tid + i * kThreads;
    if (o < np) pack(o, gq[i], cq[i]);
  }
  for (int o = tid + kPre * kThreads; o < np; o += kThreads) pack(o, g2[o], ps[o]);
  for (int o = 32; o > 0; o >>= 1) eke += __shfl_down(eke, o, 64);
  if ((tid & 63) == 0) red[tid >> 6] = eke;
  __syncthreads();
  if (tid == 0) {
    double tsum = 0.0;
    for (int w = 0; w < kThreads / 64; w++) tsum += red[w];
    c.eke_part[(size_t)s * gridDim.x + bx] = 0.5 * tsum;
  }
}

// ------------------------------------------------------------------ y passes

template <typename T, int R1, int R2>
__global__ __launch_bounds__(kThreads) void k2_yfwd(FftCtx c) {
  using Sh = Shape<R1, R2>;
  constexpr int N = Sh::N, G = Sh::G, NTR = Sh::NTR;
  __shared__ cpx<T> ex[NTR * N];
  const int kx0 = blockIdx.x * NTR, z = blockIdx.y, s = blockIdx.z;
  const int npx = c.g.nplane_x;
  const int nk = min(NTR, npx - kx0);
  const int tid = threadIdx.x;
  const size_t off = ((size_t)s * c.g.nz + z) * (size_t)npx * N;
  {  // pass 1: lanes along y of one column (contiguous in zbuf)
    const int tr = tid / G, j = tid - tr * G;
    if (tr < nk && j < R2) {
      const int kx = kx0 + tr;
      const int lo = c.kylo[kx], hi = c.kyhi[kx];
      const cpx<T>* src = (const cpx<T>*)c.zbuf + off + (size_t)kx * N;
      cpx<T> v[R1];
#pragma unroll
      for (int t = 0; t < R1; t++) {
        const int y = j + t * R2;
        const int kyv = signed_k(y, N);
        v[t] = (kyv >= lo && kyv <= hi) ? src[y] : cpx<T>{(T)0, (T)0};
      }
      Dft<T, R1, -1>::run(v);
#pragma unroll
      for (int t = 0; t < R1; t++) ex[tr * N + j * R1 + t] = v[t];
    }
  }
  __syncthreads();
  {  // pass 2: lanes along kx (contiguous in xbuf)
    const int tr = tid % NTR, j = tid / NTR;
    if (tr < nk && j < R1) {
      cpx<T> u[R2];
#pragma unroll
      for (int t = 0; t < R2; t++) u[t] = ex[tr * N + j + t * R1];
#pragma unroll
      for (int t = 1; t < R2; t++) u[t] = cm(u[t], twiddle<T, -1>(c.twy, t * j));
      Dft<T, R2, -1>::run(u);
      cpx<T>* dst = (cpx<T>*)c.xbuf + off + kx0 + tr;
#pragma unroll
      for (int t = 0; t < R2; t++) dst[(size_t)(j + t * R1) * npx] = u[t];
    }
  }
}

template <typename T, int R1, int R2>
__global__ __launch_bounds__(kThreads) void k2_ybwd(FftCtx c) {
  using Sh = Shape<R1, R2>;
  constexpr int N = Sh::N, G = Sh::G, NTR = Sh::NTR;
  __shared__ cpx<T> ex[NTR * N];
  const int kx0 = blockIdx.x * NTR, z = blockIdx.y, s = blockIdx.z;
  const int npx = c.g.nplane_x;
  const int nk = min(NTR, npx - kx0);
  const int tid = threadIdx.x;
  const size_t off = ((size_t)s * c.g.nz + z) * (size_t)npx * N;
  {  // pass 1: lanes along kx (contiguous in xbuf)
    const int tr = tid % NTR, j = tid / NTR;
    if (tr < nk && j < R2) {
      const cpx<T>* src = (const cpx<T>*)c.xbuf + off + kx0 + tr;
      cpx<T> v[R1];
#pragma unroll
      for (int t = 0; t < R1; t++) v[t] = src[(size_t)(j + t * R2) * npx];
      Dft<T, R1, +1>::run(v);
#pragma unroll
      for (int t = 0; t < R1; t++) ex[tr * N + j * R1 + t] = v[t];
    }
  }
  __syncthreads();
  {  // pass 2: lanes along y of one column
    const int tr = tid / G, j = tid - tr * G;
    if (tr < nk && j < R1) {
      cpx<T> u[R2];
#pragma unroll
      for (int t = 0; t < R2; t++) u[t] = ex[tr * N + j + t * R1];
#pragma unroll
      for (int t = 1; t < R2; t++) u[t] = cm(u[t], twiddle<T, +1>(c.twy, t * j));
      Dft<T, R2, +1>::run(u);
      const int kx = kx0 + tr;
      const int lo = c.kylo[kx], hi = c.kyhi[kx];
      cpx<T>* dst = (cpx<T>*)c.zbuf + off + (size_t)kx * N;
#pragma unroll
      for (int t = 0; t < R2; t++) {
        const int y = j + t * R1;
        const int kyv = signed_k(y, N);
        if (kyv >= lo && kyv <= hi) dst[y] = u[t];
      }
    }
  }
}

// ------------------------------------------------------------------ x passes
// Workgroup = (block of NTR pair-rows, z, chunk of states); loops over the
// chunk's states.  Lanes run along x in both passes (rows are contiguous).

template <typename T, int R1, int R2>
__global__ __launch_bounds__(kThreads) void k2_xc2r(FftCtx c) {
  using Sh = Shape<R1, R2>;
  constexpr int N = Sh::N, G = Sh::G, NTR = Sh::NTR;
  __shared__ cpx<T> ex[NTR * N];
  const int ny = c.g.ny, npx = c.g.nplane_x;
  const int npairs = (ny + 1) / 2;
  const int q0 = blockIdx.x * NTR, z = blockIdx.y, chunk = blockIdx.z;
  const int nq = min(NTR, npairs - q0);
  const int st0 = chunk * c.cs, st1 = min(c.g.nst, st0 + c.cs);
  const int tid = threadIdx.x;
  const int tr = tid / G, j = tid - tr * G;
  const int y1 = 2 * (q0 + tr), y2 = y1 + 1;
  const size_t NP = (size_t)N * ny, Ntot = NP * c.g.nz;
  double acc[2 * R2];
#pragma unroll
  for (int i = 0; i < 2 * R2; i++) acc[i] = 0.0;
  for (int s = st0; s < st1; s++) {
    if (tr < nq && j < R2) {
      const cpx<T>* xs = (const cpx<T>*)c.xbuf + ((size_t)s * c.g.nz + z) * ny * npx;
      cpx<T> v[R1];
#pragma unroll
      for (int t = 0; t < R1; t++) {
        const int k = j + t * R2;
        cpx<T> Z = {(T)0, (T)0};
        int kk = -1;
        bool mirror = false;
        if (k < npx) kk = k;
        else if (k > N - npx) { kk = N - k; mirror = true; }
        if (kk >= 0) {
          const cpx<T> X1 = xs[(size_t)y1 * npx + kk];
          const cpx<T> X2 = (y2 < ny) ? xs[(size_t)y2 * npx + kk] : cpx<T>{(T)0, (T)0};
          if (kk == 0) Z = {X1.x, X2.x};
          else if (!mirror) Z = {X1.x - X2.y, X1.y + X2.x};
          else Z = {X1.x + X2.y, X2.x - X1.y};
        }
        v[t] = Z;
      }
      Dft<T, R1, -1>::run(v);
#pragma unroll
      for (int t = 0; t < R1; t++) ex[tr * N + j * R1 + t] = v[t];
    }
    __syncthreads();
    if (tr < nq && j < R1) {
      cpx<T> u[R2];
#pragma unroll
      for (int t = 0; t < R2; t++) u[t] = ex[tr * N + j + t * R1];
#pragma unroll
      for (int t = 1; t < R2; t++) u[t] = cm(u[t], twiddle<T, -1>(c.twx, t * j));
      Dft<T, R2, -1>::run(u);
      const double oc = c.occ[s];
      T* pr = (T*)c.psir + (size_t)s * Ntot + (size_t)z * NP;
#pragma unroll
      for (int t = 0; t < R2; t++) {
        const int x = j + t * R1;
        pr[(size_t)y1 * N + x] = u[t].x;
        acc[2 * t] += oc * (double)u[t].x * (double)u[t].x;
        if (y2 < ny) {
          pr[(size_t)y2 * N + x] = u[t].y;
          acc[2 * t + 1] += oc * (double)u[t].y * (double)u[t].y;
        }
      }
    }
    __syncthreads();
  }
  if (tr < nq && j < R1) {
    double* rp = c.rho_part + (size_t)chunk * Ntot + (size_t)z * NP;
#pragma unroll
    for (int t = 0; t < R2; t++) {
      const int x = j + t * R1;
      rp[(size_t)y1 * N + x] = acc[2 * t];
      if (y2 < ny) rp[(size_t)y2 * N + x] = acc[2 * t + 1];
    }
  }
}

template <typename T, int R1, int R2>
__global__ __launch_bounds__(kThreads) void k2_xr2c(FftCtx c) {
  using Sh = Shape<R1, R2>;
  constexpr int N = Sh::N, G = Sh::G, NTR = Sh::NTR;
  __shared__ cpx<T> ex[NTR * N];
  const int ny = c.g.ny, npx = c.g.nplane_x;
  const int npairs = (ny + 1) / 2;
  const int q0 = blockIdx.x * NTR, z = blockIdx.y, chunk = blockIdx.z;
  const int nq = min(NTR, npairs - q0);
  const int st0 = chunk * c.cs, st1 = min(c.g.nst, st0 + c.cs);
  const int tid = threadIdx.x;
  const int tr = tid / G, j = tid - tr * G;
  const int y1 = 2 * (q0 + tr), y2 = y1 + 1;
  const size_t NP = (size_t)N * ny, Ntot = NP * c.g.nz;
  // V at this thread's pass-1 positions, both rows
  T v1[R1], v2[R1];
#pragma unroll
  for (int t = 0; t < R1; t++) {
    v1[t] = (T)0;
    v2[t] = (T)0;
    if (tr < nq && j < R2) {
      const double* vz = c.vks + (size_t)z * NP;
      const int x = j + t * R2;
      v1[t] = (T)vz[(size_t)y1 * N + x];
      if (y2 < ny) v2[t] = (T)vz[(size_t)y2 * N + x];
    }
  }
  for (int s = st0; s < st1; s++) {
    if (tr < nq && j < R2) {
      const T oc = (T)c.occ[s];
      const T* pr = (const T*)c.psir + (size_t)s * Ntot + (size_t)z * NP;
      cpx<T> v[R1];
#pragma unroll
      for (int t = 0; t < R1; t++) {
        const int x = j + t * R2;
        const T f1 = pr[(size_t)y1 * N + x] * (v1[t] * oc);
        const T f2 = (y2 < ny) ? pr[(size_t)y2 * N + x] * (v2[t] * oc) : (T)0;
        v[t] = {f1, f2};
      }
      Dft<T, R1, -1>::run(v);
#pragma unroll
      for (int t = 0; t < R1; t++) ex[tr * N + j * R1 + t] = v[t];
    }
    __syncthreads();
    if (tr < nq && j < R1) {
      cpx<T> u[R2];
#pragma unroll
      for (int t = 0; t < R2; t++) u[t] = ex[tr * N + j + t * R1];
#pragma unroll
      for (int t = 1; t < R2; t++) u[t] = cm(u[t], twiddle<T, -1>(c.twx, t * j));
      Dft<T, R2, -1>::run(u);
#pragma unroll
      for (int t = 0; t < R2; t++) ex[tr * N + j + t * R1] = u[t];
    }
    __syncthreads();
    // split the two real rows and keep kx < nplane_x, conjugated
    cpx<T>* xs = (cpx<T>*)c.xbuf + ((size_t)s * c.g.nz + z) * ny * npx;
    for (int idx = tid; idx < nq * npx; idx += kThreads) {
      const int q = idx / npx, k = idx - q * npx;
      const int r1 = 2 * (q0 + q), r2 = r1 + 1;
      const cpx<T> Zk = ex[q * N + k];
      const cpx<T> Zm = ex[q * N + (k == 0 ? 0 : N - k)];
      xs[(size_t)r1 * npx + k] = {(T)0.5 * (Zk.x + Zm.x), -(T)0.5 * (Zk.y - Zm.y)};
      if (r2 < ny) xs[(size_t)r2 * npx + k] = {(T)0.5 * (Zk.y + Zm.y), (T)0.5 * (Zk.x - Zm.x)};
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ launch

enum Kind { ZFWD = 0, YFWD, XC2R, XR2C, YBWD, ZBWD };

template <typename T, int R1, int R2>
hipError_t launch(int kind, const FftCtx& c, hipStream_t st) {
  constexpr int NTR = Shape<R1, R2>::NTR;
  const Geom& g = c.g;
  switch (kind) {
    case ZFWD:
      if (c.prune_z)
        hipLaunchKernelGGL((k2_zfwd<T, R1, R2, true>), dim3((g.nlines + NTR - 1) / NTR, g.nst),
                           dim3(kThreads), 0, st, c);
      else
        hipLaunchKernelGGL((k2_zfwd<T, R1, R2, false>), dim3((g.nlines + NTR - 1) / NTR, g.nst),
                           dim3(kThreads), 0, st, c);
      break;
    case ZBWD:
      if (c.prune_z)
        hipLaunchKernelGGL((k2_zbwd<T, R1, R2, true>), dim3((g.nlines + NTR - 1) / NTR, g.nst),
                           dim3(kThreads), 0, st, c);
      else
        hipLaunchKernelGGL((k2_zbwd<T, R1, R2, false>), dim3((g.nlines + NTR - 1) / NTR, g.nst),
                           dim3(kThreads), 0, st, c);
      break;
    case YFWD:
      hipLaunchKernelGGL((k2_yfwd<T, R1, R2>), dim3((g.nplane_x + NTR - 1) / NTR, g.nz, g.nst),
                         dim3(kThreads), 0, st, c);
      break;
    case YBWD:
      hipLaunchKernelGGL((k2_ybwd<T, R1, R2>), dim3((g.nplane_x + NTR - 1) / NTR, g.nz, g.nst),
                         dim3(kThreads), 0, st, c);
      break;
    case XC2R:
      hipLaunchKernelGGL((k2_xc2r<T, R1, R2>),
                         dim3(((g.ny + 1) / 2 + NTR - 1) / NTR, g.nz, c.nchunk), dim3(kThreads),
                         0, st, c);
      break;
    case XR2C:
      hipLaunchKernelGGL((k2_xr2c<T, R1, R2>),
                         dim3(((g.ny + 1) / 2 + NTR - 1) / NTR, g.nz, c.nchunk), dim3(kThreads),
                         0, st, c);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace oa2
