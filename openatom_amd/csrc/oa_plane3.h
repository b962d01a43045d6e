// oa_plane3.h -- fused y+x plane passes of the state FFT chain on the
// two-factor register engine (oa_fft2.h).
//
// One workgroup owns one z-plane and loops over a chunk of states.  Per state
// it reads the populated columns of the z-transformed plane (zbuf), runs the
// y-FFT of the nplane_x columns into an LDS plane buffer P[kx][y], then the
// x passes in rounds of pair-rows (two real rows per complex transform) with
// an LDS exchange buffer E.  Nothing of psi(r) goes to HBM:
//
//   k3_g2r : y-FFT (FFTW_FORWARD) + Hermitian x-FFT -> rho partials of the
//            chunk, accumulated in registers over the chunk's states
//            (CP_State_RealSpacePlane.C:349-355, 450-517);
//   k3_r2g : the same G->R recomputed from zbuf, psi(r) V(r) occ, x-FFT
//            (r2c, conjugated) and y-FFT (FFTW_BACKWARD) -> zout
//            (CP_State_RealSpacePlane.C:530-597, fftCache.C:1252-1319).
//
// Recomputing the G->R plane in k3_r2g costs one y-FFT and one x-FFT per
// plane; it replaces a write and a read of psi(r) (Nx*Ny*Nz reals per state)
// with a second read of the populated z-lines.  rho is complete before V
// psi is formed, as in the reference, so V may depend on rho.
//
// Template parameters: transform lengths Nx = X1*X2 and Ny = Y1*Y2 (the x
// pass-2 writes X2 outputs per thread, so the smaller factor goes second),
// NT threads.  P column stride CY = Ny | 1.
#pragma once

#include "oa_fft2.h"
#include "oa_kernels.h"

namespace oa3 {

using oa::FftCtx;
using oa2::cm;
using oa2::Dft;
using oa2::twiddle;

__device__ __forceinline__ int signed_k(int i, int n) { return (i <= (n - 1) / 2) ? i : i - n; }

template <int X1, int X2, int Y1, int Y2, int NT>
struct PS {
  static constexpr int NX = X1 * X2, NY = Y1 * Y2;
  static constexpr int GX = X1 > X2 ? X1 : X2;  // threads per x transform
  static constexpr int GY = Y1 > Y2 ? Y1 : Y2;  // threads per y transform
  static constexpr int NTRX = NT / GX;          // pair-rows per x round
  static constexpr int NTRY = NT / GY;          // columns per y batch
  static constexpr int NPAIR = (NY + 1) / 2;
  static constexpr int ROUNDS = (NPAIR + NTRX - 1) / NTRX;
  static constexpr int CY = NY | 1;
  static constexpr int EX = NTRX * NX > NTRY * NY ? NTRX * NX : NTRY * NY;  // E slots
  // pass-2 twiddles W_N^m, m = t j <= (X1 - 1)(X2 - 1), in LDS (planes are
  // square, so one table serves the x and the y transforms)
  static constexpr int NTW = (X1 - 1) * (X2 - 1) + 1;
  static_assert(NX == NY, "the fused plane passes need square x-y planes");
};

// pass-2 twiddle W_N^m from the LDS table, conjugated for S = +1
template <typename T, int S>
__device__ __forceinline__ cpx<T> twl(const cpx<T>* tw, int m) {
  const cpx<T> w = tw[m];
  return {w.x, S < 0 ? w.y : -w.y};
}

// copy the pass-2 twiddles and the per-column ky ranges into LDS (callers
// synchronise before use); the ranges are read for every state, and a
// global load there would cost a memory latency per state
template <typename T, int NTW, int NT>
__device__ __forceinline__ void load_twl(cpx<T>* tw, const double2* __restrict__ g, int2* ky,
                                         const FftCtx& c, int npx) {
  for (int i = threadIdx.x; i < NTW; i += NT) tw[i] = {(T)g[i].x, (T)g[i].y};
  for (int i = threadIdx.x; i < npx; i += NT) ky[i] = make_int2(c.kylo[i], c.kyhi[i]);
}

// LDS bytes of either kernel: P (nplane_x columns) + E + twiddles + ky ranges
template <typename T, int X1, int X2, int Y1, int Y2, int NT>
size_t plane3_smem(int npx) {
  using S = PS<X1, X2, Y1, Y2, NT>;
  return sizeof(cpx<T>) * ((size_t)npx * S::CY + S::EX + S::NTW) + sizeof(int2) * npx;
}

// y-FFT (FFTW_FORWARD) of the populated columns of one z-plane into P.
// zs: zbuf plane [kx][NY]; only ky in KY[kx] = [kylo, kyhi] of each column
// is read.
template <typename T, int Y1, int Y2, int NT>
__device__ __forceinline__ void y_fwd(const cpx<T>* __restrict__ zs, cpx<T>* P, const cpx<T>* TW,
                                      const int2* KY, int npx) {
  constexpr int NY = Y1 * Y2, GY = Y1 > Y2 ? Y1 : Y2, NTRY = NT / GY, CY = NY | 1;
  const int tid = threadIdx.x;
  const int tr = tid / GY, j = tid - tr * GY;
  for (int cb = 0; cb < npx; cb += NTRY) {
    const int kx = cb + tr;
    if (tr < NTRY && kx < npx && j < Y2) {
      const int lo = KY[kx].x, hi = KY[kx].y;
      const cpx<T>* src = zs + (size_t)kx * NY;
      cpx<T> v[Y1];
#pragma unroll
      for (int t = 0; t < Y1; t++) {
        const int y = j + t * Y2;
        const int kyv = signed_k(y, NY);
        v[t] = (kyv >= lo && kyv <= hi) ? src[y] : cpx<T>{(T)0, (T)0};
      }
      Dft<T, Y1, -1>::run(v);
#pragma unroll
      for (int t = 0; t < Y1; t++) P[kx * CY + j * Y1 + t] = v[t];
    }
  }
  __syncthreads();
  // pass 2 in place: thread j reads and writes the slots j + t*Y1 of its column
  for (int cb = 0; cb < npx; cb += NTRY) {
    const int kx = cb + tr;
    if (tr < NTRY && kx < npx && j < Y1) {
      cpx<T> u[Y2];
#pragma unroll
      for (int t = 0; t < Y2; t++) u[t] = P[kx * CY + j + t * Y1];
#pragma unroll
      for (int t = 1; t < Y2; t++) u[t] = cm(u[t], twl<T, -1>(TW, t * j));
      Dft<T, Y2, -1>::run(u);
#pragma unroll
      for (int t = 0; t < Y2; t++) P[kx * CY + j + t * Y1] = u[t];
    }
  }
  __syncthreads();
}

// y-FFT (FFTW_BACKWARD) of the nplane_x columns of P into the zout plane,
// keeping ky in KY[kx]; E is the exchange buffer.
template <typename T, int Y1, int Y2, int NT>
__device__ __forceinline__ void y_bwd(const cpx<T>* P, cpx<T>* E, const cpx<T>* TW,
                                      cpx<T>* __restrict__ zo, const int2* KY, int npx) {
  constexpr int NY = Y1 * Y2, GY = Y1 > Y2 ? Y1 : Y2, NTRY = NT / GY, CY = NY | 1;
  const int tid = threadIdx.x;
  const int tr = tid / GY, j = tid - tr * GY;
  for (int cb = 0; cb < npx; cb += NTRY) {
    const int kx = cb + tr;
    const bool act = tr < NTRY && kx < npx;
    if (act && j < Y2) {
      cpx<T> v[Y1];
#pragma unroll
      for (int t = 0; t < Y1; t++) v[t] = P[kx * CY + j + t * Y2];
      Dft<T, Y1, +1>::run(v);
#pragma unroll
      for (int t = 0; t < Y1; t++) E[tr * NY + j * Y1 + t] = v[t];
    }
    __syncthreads();
    if (act && j < Y1) {
      cpx<T> u[Y2];
#pragma unroll
      for (int t = 0; t < Y2; t++) u[t] = E[tr * NY + j + t * Y1];
#pragma unroll
      for (int t = 1; t < Y2; t++) u[t] = cm(u[t], twl<T, +1>(TW, t * j));
      Dft<T, Y2, +1>::run(u);
      const int lo = KY[kx].x, hi = KY[kx].y;
      cpx<T>* dst = zo + (size_t)kx * NY;
#pragma unroll
      for (int t = 0; t < Y2; t++) {
        const int y = j + t * Y1;
        const int kyv = signed_k(y, NY);
        if (kyv >= lo && kyv <= hi) dst[y] = u[t];
      }
    }
    __syncthreads();
  }
}

// One round of Hermitian x-FFTs (c2r by the pair-row trick) from the rows of
// P: pair q carries rows y1 = 2q (real part) and y2 = 2q+1 (imaginary part).
// On return thread (tr, j < X1) holds psi(r) at x = j + t*X1 in u[t].
// Ends after the pass-2 reads of E (callers sync before reusing E).
template <typename T, int X1, int X2, int NT>
__device__ __forceinline__ void x_c2r(const cpx<T>* P, cpx<T>* E, const cpx<T>* TW,
                                      const FftCtx& c, int npx, int ny, int q, int tr, int j,
                                      bool act, cpx<T> (&u)[X2]) {
  constexpr int NX = X1 * X2;
  const int CY = ny | 1;
  const int y1 = 2 * q, y2 = y1 + 1;
  if (act && j < X2) {
    cpx<T> v[X1];
#pragma unroll
    for (int t = 0; t < X1; t++) {
      const int k = j + t * X2;
      cpx<T> Z = {(T)0, (T)0};
      int kk = -1;
      bool mirror = false;
      if (k < npx) {
        kk = k;
      } else if (k > NX - npx) {
        kk = NX - k;
        mirror = true;
      }
      if (kk >= 0) {
        const cpx<T> A = P[kk * CY + y1];
        const cpx<T> B = (y2 < ny) ? P[kk * CY + y2] : cpx<T>{(T)0, (T)0};
        if (kk == 0) Z = {A.x, B.x};
        else if (!mirror) Z = {A.x - B.y, A.y + B.x};
        else Z = {A.x + B.y, B.x - A.y};
      }
      v[t] = Z;
    }
    Dft<T, X1, -1>::run(v);
#pragma unroll
    for (int t = 0; t < X1; t++) E[tr * NX + j * X1 + t] = v[t];
  }
  __syncthreads();
  if (act && j < X1) {
#pragma unroll
    for (int t = 0; t < X2; t++) u[t] = E[tr * NX + j + t * X1];
#pragma unroll
    for (int t = 1; t < X2; t++) u[t] = cm(u[t], twl<T, -1>(TW, t * j));
    Dft<T, X2, -1>::run(u);
  }
}

// G->R plane pass.  Grid (nz, row parts, state chunks).  Row part rp covers
// x rounds [rp*RPW, rp*RPW + RPW).  PSIR: write psi(r) of state
// c.psir_state into c.psir ([Nz][Ny][Nx] of T) instead of rho partials.
template <typename T, int X1, int X2, int Y1, int Y2, int NT, int RPW, bool PSIR>
__global__ __launch_bounds__(NT) void k3_g2r(FftCtx c) {
  using S = PS<X1, X2, Y1, Y2, NT>;
  constexpr int NX = S::NX, NY = S::NY, GX = S::GX, NTRX = S::NTRX;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int npx = c.g.nplane_x;
  cpx<T>* P = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* E = P + (size_t)npx * S::CY;
  cpx<T>* TW = E + S::EX;
  int2* KY = reinterpret_cast<int2*>(TW + S::NTW);
  const int z = blockIdx.x, rp = blockIdx.y, chunk = blockIdx.z;
  const int r0 = rp * RPW;
  int st0, st1;
  if (PSIR) {
    st0 = c.psir_state;
    st1 = st0 + 1;
  } else {
    st0 = chunk * c.cs;
    st1 = min(c.g.nst, st0 + c.cs);
  }
  const int tid = threadIdx.x;
  const int tr = tid / GX, j = tid - tr * GX;
  const size_t plane = (size_t)npx * NY;
  const size_t NP = (size_t)NX * NY;
  load_twl<T, S::NTW, NT>(TW, c.twx, KY, c, npx);
  __syncthreads();
  double acc[RPW][2 * X2];
#pragma unroll
  for (int a = 0; a < RPW; a++)
#pragma unroll
    for (int b = 0; b < 2 * X2; b++) acc[a][b] = 0.0;
  for (int s = st0; s < st1; s++) {
    const cpx<T>* zs = (const cpx<T>*)c.zbuf + ((size_t)s * c.g.nz + z) * plane;
    y_fwd<T, Y1, Y2, NT>(zs, P, TW, KY, npx);
    const double oc = c.occ[s];
#pragma unroll
    for (int rr = 0; rr < RPW; rr++) {
      const int r = r0 + rr;
      if (r < S::ROUNDS) {
        const int q = r * NTRX + tr;
        const bool act = tr < NTRX && q < S::NPAIR;
        cpx<T> u[X2];
        x_c2r<T, X1, X2, NT>(P, E, TW, c, npx, NY, q, tr, j, act, u);
        if (act && j < X1) {
          const int y1 = 2 * q, y2 = y1 + 1;
          if (PSIR) {
            T* pr = (T*)c.psir + (size_t)z * NP;
#pragma unroll
            for (int t = 0; t < X2; t++) {
              pr[(size_t)y1 * NX + j + t * X1] = u[t].x;
              if (y2 < NY) pr[(size_t)y2 * NX + j + t * X1] = u[t].y;
            }
          } else {
#pragma unroll
            for (int t = 0; t < X2; t++) {
              acc[rr][2 * t] += oc * (double)u[t].x * (double)u[t].x;
              acc[rr][2 * t + 1] += oc * (double)u[t].y * (double)u[t].y;
            }
          }
        }
        __syncthreads();
      }
    }
  }
  if (PSIR) return;
  double* rp_out = c.rho_part + (size_t)chunk * NP * c.g.nz + (size_t)z * NP;
#pragma unroll
  for (int rr = 0; rr < RPW; rr++) {
    const int r = r0 + rr;
    const int q = r * NTRX + tr;
    if (r < S::ROUNDS && tr < NTRX && q < S::NPAIR && j < X1) {
      const int y1 = 2 * q, y2 = y1 + 1;
#pragma unroll
      for (int t = 0; t < X2; t++) {
        rp_out[(size_t)y1 * NX + j + t * X1] = acc[rr][2 * t];
        if (y2 < NY) rp_out[(size_t)y2 * NX + j + t * X1] = acc[rr][2 * t + 1];
      }
    }
  }
}

// R->G plane pass.  Grid (nz, state chunks).  VREG: V(r) of the thread's
// pass-2 positions for all rounds is held in registers across the chunk.
// FP32 up to N = 100: at most 168 VGPRs, three waves per SIMD (measured
// faster than two; larger planes would spill).  RHO: also accumulate rho
// partials of the chunk from the recomputed psi(r), for a V(r) that does
// not depend on this step's rho (then no separate G->R pass is needed).
template <typename T, int X1, int X2, int Y1, int Y2, int NT, bool VREG, bool RHO>
__global__ __launch_bounds__(NT, (sizeof(T) == 4 && !RHO) ? (X1 * X2 <= 100 ? 3 : 2) : 1) void
k3_r2g(FftCtx c) {
  using S = PS<X1, X2, Y1, Y2, NT>;
  constexpr int NX = S::NX, NY = S::NY, GX = S::GX, NTRX = S::NTRX, R = S::ROUNDS;
  constexpr int CY = S::CY;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int npx = c.g.nplane_x;
  cpx<T>* P = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* E = P + (size_t)npx * CY;
  cpx<T>* TW = E + S::EX;
  int2* KY = reinterpret_cast<int2*>(TW + S::NTW);
  const int z = blockIdx.x, chunk = blockIdx.y;
  const int st0 = chunk * c.cs_r, st1 = min(c.g.nst, st0 + c.cs_r);
  const int tid = threadIdx.x;
  const int tr = tid / GX, j = tid - tr * GX;
  const size_t plane = (size_t)npx * NY;
  const size_t NP = (size_t)NX * NY;
  const double* vz = c.vks + (size_t)z * NP;
  load_twl<T, S::NTW, NT>(TW, c.twx, KY, c, npx);
  __syncthreads();
  T vr[VREG ? R : 1][VREG ? 2 * X2 : 1];
  double acc[RHO ? R : 1][RHO ? 2 * X2 : 1];
  if (RHO) {
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int b = 0; b < 2 * X2; b++) acc[r][b] = 0.0;
  }
  // V loads are branch-free (clamped rows and columns, masked after the
  // load): a load inside a divergent branch is waited for on its own, so a
  // series of them costs one memory latency each
  if (VREG) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int q = r * NTRX + tr;
      const bool act = tr < NTRX && q < S::NPAIR && j < X1;
      const int ya = 2 * min(q, S::NPAIR - 1), yb = min(ya + 1, NY - 1);
#pragma unroll
      for (int t = 0; t < X2; t++) {
        const int x = min(j + t * X1, NX - 1);
        const T a = (T)vz[(size_t)ya * NX + x], b = (T)vz[(size_t)yb * NX + x];
        vr[r][2 * t] = act ? a : (T)0;
        vr[r][2 * t + 1] = (act && 2 * q + 1 < NY) ? b : (T)0;
      }
    }
  }
  for (int s = st0; s < st1; s++) {
    const cpx<T>* zs = (const cpx<T>*)c.zbuf + ((size_t)s * c.g.nz + z) * plane;
    y_fwd<T, Y1, Y2, NT>(zs, P, TW, KY, npx);
    const T oc = (T)c.occ[s];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int q = r * NTRX + tr;
      const bool act = tr < NTRX && q < S::NPAIR;
      // without VREG: this round's V, loaded (branch-free) ahead of the transform
      T vv[VREG ? 1 : 2 * X2];
      if (!VREG) {
        const int ya = 2 * min(q, S::NPAIR - 1), yb = min(ya + 1, NY - 1);
#pragma unroll
        for (int t = 0; t < X2; t++) {
          const int x = min(j + t * X1, NX - 1);
          vv[VREG ? 0 : 2 * t] = (T)vz[(size_t)ya * NX + x];
          vv[VREG ? 0 : 2 * t + 1] = (T)vz[(size_t)yb * NX + x];
        }
      }
      cpx<T> u[X2];
      x_c2r<T, X1, X2, NT>(P, E, TW, c, npx, NY, q, tr, j, act, u);
      __syncthreads();  // E is rewritten below
      if (act && j < X1) {
        const int y1 = 2 * q, y2 = y1 + 1;
        if (RHO) {
          const double od = c.occ[s];
#pragma unroll
          for (int t = 0; t < X2; t++) {
            acc[r][2 * t] += od * (double)u[t].x * (double)u[t].x;
            acc[r][2 * t + 1] += od * (double)u[t].y * (double)u[t].y;
          }
        }
        cpx<T> w[X2];
#pragma unroll
        for (int t = 0; t < X2; t++) {
          T v1, v2;
          if (VREG) {
            v1 = vr[r][2 * t];
            v2 = vr[r][2 * t + 1];
          } else {
            v1 = vv[VREG ? 0 : 2 * t];
            v2 = vv[VREG ? 0 : 2 * t + 1];
          }
          w[t] = {u[t].x * (v1 * oc), (y2 < NY) ? u[t].y * (v2 * oc) : (T)0};
        }
        // forward transform of the pair, N = X2 * X1: thread j holds
        // x = j + t*X1 (t < X2), so pass 1 is a radix-X2 DFT in registers
        Dft<T, X2, -1>::run(w);
#pragma unroll
        for (int t = 0; t < X2; t++) E[tr * NX + j * X2 + t] = w[t];
      }
      __syncthreads();
      if (act && j < X2) {
        cpx<T> a[X1];
#pragma unroll
        for (int t = 0; t < X1; t++) a[t] = E[tr * NX + j + t * X2];
#pragma unroll
        for (int t = 1; t < X1; t++) a[t] = cm(a[t], twl<T, -1>(TW, t * j));
        Dft<T, X1, -1>::run(a);
#pragma unroll
        for (int t = 0; t < X1; t++) E[tr * NX + j + t * X2] = a[t];
      }
      __syncthreads();
      // split the two real rows, keep kx < nplane_x, conjugated, into P
      const int nq = min(NTRX, S::NPAIR - r * NTRX);
      for (int idx = tid; idx < nq * npx; idx += NT) {
        const int qq = idx / npx, k = idx - qq * npx;
        const int r1 = 2 * (r * NTRX + qq), r2 = r1 + 1;
        const cpx<T> Zk = E[qq * NX + k];
        const cpx<T> Zm = E[qq * NX + (k == 0 ? 0 : NX - k)];
        P[k * CY + r1] = {(T)0.5 * (Zk.x + Zm.x), -(T)0.5 * (Zk.y - Zm.y)};
        if (r2 < NY) P[k * CY + r2] = {(T)0.5 * (Zk.y + Zm.y), (T)0.5 * (Zk.x - Zm.x)};
      }
      __syncthreads();
    }
    cpx<T>* zo = (cpx<T>*)c.zout + ((size_t)s * c.g.nz + z) * plane;
    y_bwd<T, Y1, Y2, NT>(P, E, TW, zo, KY, npx);
  }
  if (!RHO) return;
  // rho partial of this chunk (same layout and per-element order as k3_g2r)
  double* rp_out = c.rho_part + (size_t)chunk * NP * c.g.nz + (size_t)z * NP;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int q = r * NTRX + tr;
    if (tr < NTRX && q < S::NPAIR && j < X1) {
      const int y1 = 2 * q, y2 = y1 + 1;
#pragma unroll
      for (int t = 0; t < X2; t++) {
        rp_out[(size_t)y1 * NX + j + t * X1] = acc[r][2 * t];
        if (y2 < NY) rp_out[(size_t)y2 * NX + j + t * X1] = acc[r][2 * t + 1];
      }
    }
  }
}

// rho accumulation fits the R->G pass's registers: at most two rounds of
// radix-10 pair rows
template <int X1, int X2, int NT>
constexpr bool r2g_rho_ok() {
  return PS<X1, X2, X2, X1, NT>::ROUNDS * X2 <= 20;
}

// ------------------------------------------------------------------ launch

enum PlaneKind { G2R = 0, R2G = 1, G2R_PSIR = 2, R2G_RHO = 3 };

constexpr int kPlaneThreads = 256;

// x rounds per k3_g2r workgroup: the rho accumulators are 2*X2 doubles per
// round and thread; FP64 keeps one round, FP32 two.  When the LDS of one
// workgroup exceeds half a CU's (one workgroup per CU, four waves), the
// register file has room for every round and one workgroup covers the whole
// plane ("wide"), so the y-FFT of a plane is not repeated per row part.
template <typename T>
constexpr int row_rounds() {
  return sizeof(T) == 8 ? 1 : 2;
}

constexpr size_t kCuLds = 160 * 1024;

template <typename T, int X1, int X2, int Y1, int Y2, int NT>
bool g2r_wide(int npx) {
  return 2 * plane3_smem<T, X1, X2, Y1, Y2, NT>(npx) > kCuLds;
}

// Launch facts of one size for the host's chunk choice: row parts of the
// G->R grid and resident workgroups per CU of both kernels.
template <typename T, int X1, int X2, int Y1, int Y2>
hipError_t plane3_info(int npx, int* rowparts, int* occ_g2r, int* occ_r2g, int* rho_ok) {
  constexpr int NT = kPlaneThreads;
  using S = PS<X1, X2, Y1, Y2, NT>;
  constexpr bool VREG = sizeof(T) == 4 && S::ROUNDS * 2 * X2 <= 48;
  const size_t smem = plane3_smem<T, X1, X2, Y1, Y2, NT>(npx);
  const bool wide = g2r_wide<T, X1, X2, Y1, Y2, NT>(npx);
  *rowparts = wide ? 1 : (S::ROUNDS + row_rounds<T>() - 1) / row_rounds<T>();
  const void* g = wide ? (const void*)k3_g2r<T, X1, X2, Y1, Y2, NT, S::ROUNDS, false>
                       : (const void*)k3_g2r<T, X1, X2, Y1, Y2, NT, row_rounds<T>(), false>;
  const void* r = (const void*)k3_r2g<T, X1, X2, Y1, Y2, NT, VREG, false>;
  *rho_ok = r2g_rho_ok<X1, X2, NT>() ? 1 : 0;
  hipError_t e;
  if ((e = oa::grant_lds(g, smem)) != hipSuccess || (e = oa::grant_lds(r, smem)) != hipSuccess)
    return e;
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(occ_g2r, g, NT, smem)) != hipSuccess)
    return e;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ_r2g, r, NT, smem);
}

template <typename T, int X1, int X2, int Y1, int Y2>
hipError_t launch_plane3(int kind, const FftCtx& c, hipStream_t st) {
  constexpr int NT = kPlaneThreads;
  using S = PS<X1, X2, Y1, Y2, NT>;
  constexpr int kRowRounds = row_rounds<T>();
  constexpr int RPn = (S::ROUNDS + kRowRounds - 1) / kRowRounds;
  constexpr bool VREG = sizeof(T) == 4 && S::ROUNDS * 2 * X2 <= 48;
  const size_t smem = plane3_smem<T, X1, X2, Y1, Y2, NT>(c.g.nplane_x);
  if (c.g.nx != S::NX || c.g.ny != S::NY) return hipErrorInvalidValue;
  // raise the dynamic LDS limit of an instantiation (once per device)
  auto grant = [&](const void* fn) -> hipError_t { return oa::grant_lds(fn, smem); };
  hipError_t e = hipSuccess;
  switch (kind) {
    case G2R: {
      if (g2r_wide<T, X1, X2, Y1, Y2, NT>(c.g.nplane_x)) {
        auto fn = k3_g2r<T, X1, X2, Y1, Y2, NT, S::ROUNDS, false>;
        if ((e = grant((const void*)fn)) != hipSuccess) return e;
        hipLaunchKernelGGL(fn, dim3(c.g.nz, 1, c.nchunk), dim3(NT), smem, st, c);
      } else {
        auto fn = k3_g2r<T, X1, X2, Y1, Y2, NT, kRowRounds, false>;
        if ((e = grant((const void*)fn)) != hipSuccess) return e;
        hipLaunchKernelGGL(fn, dim3(c.g.nz, RPn, c.nchunk), dim3(NT), smem, st, c);
      }
      break;
    }
    case G2R_PSIR: {
      auto fn = k3_g2r<T, X1, X2, Y1, Y2, NT, kRowRounds, true>;
      if ((e = grant((const void*)fn)) != hipSuccess) return e;
      hipLaunchKernelGGL(fn, dim3(c.g.nz, RPn, 1), dim3(NT), smem, st, c);
      break;
    }
    case R2G: {
      auto fn = k3_r2g<T, X1, X2, Y1, Y2, NT, VREG, false>;
      if ((e = grant((const void*)fn)) != hipSuccess) return e;
      hipLaunchKernelGGL(fn, dim3(c.g.nz, c.nchunk_r), dim3(NT), smem, st, c);
      break;
    }
    case R2G_RHO: {
      if (!r2g_rho_ok<X1, X2, NT>()) return hipErrorInvalidValue;
      auto fn = k3_r2g<T, X1, X2, Y1, Y2, NT, VREG, r2g_rho_ok<X1, X2, NT>()>;
      if ((e = grant((const void*)fn)) != hipSuccess) return e;
      hipLaunchKernelGGL(fn, dim3(c.g.nz, c.nchunk_r), dim3(NT), smem, st, c);
      break;
    }
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace oa3
