// oa_plane3.h -- fused y+x plane passes of the state FFT chain on the
// two-factor register engine (oa_fft2.h).
//
// One workgroup owns one z-plane and loops over a chunk of states.  Per state
// it reads the populated columns of the z-transformed plane (zbuf), runs the
// y-FFT of the nplane_x columns into an LDS plane buffer P[kx][y], then the
// x passes in rounds of pair-rows (two real rows per complex transform) with
// an LDS exchange buffer E.  By default nothing of psi(r) goes to HBM:
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
// with a second read of the populated z-lines.  The KEEP variants take the
// other side of that trade: the G->R pass also writes psi(r) (c.psk, in the
// x pass's lane order) and the R->G pass reads it back instead of repeating
// the two transforms.  rho is complete before V psi is formed, as in the
// reference, so V may depend on rho.
//
// Template parameters: transform lengths Nx = X1*X2 and Ny = Y1*Y2 (the x
// pass-2 writes X2 outputs per thread, so the smaller factor goes second),
// NT threads, and the LDS layout L (below).
//
// LDS banks.  gfx950 LDS has 32 banks of 4 bytes: an 8-byte access (complex
// float) is served 16 lanes per cycle, a 16-byte access (complex double) 8
// lanes per cycle, and lanes of one group that hit the same bank at different
// addresses are serialised (measured: element strides 2^k cost min(2^k, 16)
// cycles, odd strides one).  A transform's G = max(R1, R2) threads do not
// fill a lane group, so a group mixes two or three transforms.  Three choices
// keep the accesses apart:
//   - the pass-2 twiddles W_N^{t j} come from tables indexed [t][j], so the
//     lanes of one transform read consecutive entries (the old table indexed
//     by t*j put them t entries apart);
//   - each phase picks its lane map: j-minor (thread = transform * G + j)
//     where its global loads or stores must run along y, transform-minor
//     (thread = j * NTR + transform) where only LDS is touched, so that the
//     lanes of a group are different transforms at the same j and differ by
//     the row or column stride;
//   - P columns and E rows have odd strides (in elements), and pass-1 output
//     blocks may be padded.
#pragma once

#include "oa_fft2.h"
#include "oa_kernels.h"

namespace oa3 {

using oa::FftCtx;
using oa2::cm;
using oa2::Dft;
using oa2::twiddle;

__device__ __forceinline__ int signed_k(int i, int n) { return (i <= (n - 1) / 2) ? i : i - n; }

// LDS layout and lane maps of the plane passes.
//   CYP : elements added to each P column beyond its padded length
//   YA  : P columns hold YA empty elements after every Y1 rows
//   EA  : exchange rows hold EA empty elements after every pass-1 output block
//   REP : elements added to each exchange row beyond its padded length
//   MAPS: transform-minor lane maps, bits: 1 y_fwd pass 2, 2 x c2r pass 1,
//         4 x c2r pass 2 (and what follows it in registers: rho, V psi and
//         the r2c pass 1), 8 r2c pass 2, 16 y_bwd pass 1
//   TW2 : twiddle tables indexed [t][j] (else one table indexed t*j)
//   PR  : prune the plane transforms to the sphere (Sph, oa_fft2.h): every
//         populated kx and ky satisfies |k| <= (N - 1) / 4
template <int CYP_, int YA_, int EA_, int REP_, int MAPS_, bool TW2_, bool PR_ = false>
struct Lay {
  static constexpr int CYP = CYP_, YA = YA_, EA = EA_, REP = REP_, MAPS = MAPS_;
  static constexpr bool TW2 = TW2_, PR = PR_;
};

using oa2::Sph;  // the sphere windows of the pruned passes (oa_fft2.h)

enum : int { kMapYf2 = 1, kMapXc1 = 2, kMapXc2 = 4, kMapRx2 = 8, kMapYb1 = 16 };

template <int X1, int X2, int Y1, int Y2, int NT, class L>
struct PS {
  static constexpr int NX = X1 * X2, NY = Y1 * Y2;
  static constexpr int GX = X1 > X2 ? X1 : X2;  // threads per x transform
  static constexpr int GY = Y1 > Y2 ? Y1 : Y2;  // threads per y transform
  static constexpr int NTRX = NT / GX;          // pair-rows per x round
  static constexpr int NTRY = NT / GY;          // columns per y batch
  static constexpr int NPAIR = (NY + 1) / 2;
  static constexpr int ROUNDS = (NPAIR + NTRX - 1) / NTRX;
  static constexpr int XMIN = X1 < X2 ? X1 : X2;
  // P column stride and exchange row strides, in elements
  static constexpr int CY = NY + L::YA * ((NY - 1) / Y1) + L::CYP;
  static constexpr int RE = NX + L::EA * ((NX - 1) / XMIN) + L::REP;
  static constexpr int RY = NY + L::EA * ((NY - 1) / Y1) + L::REP;
  static constexpr int EX = NTRX * RE > NTRY * RY ? NTRX * RE : NTRY * RY;  // E slots
  // twiddles: [t][j] tables A (t < X1 = Y2, j < X2 = Y1) and B (t < X2,
  // j < X1; the same table when X1 == X2), or one table by m = t j
  static constexpr int NTWA = X1 * X2;
  static constexpr int NTWB = X1 == X2 ? 0 : X1 * X2;
  static constexpr int NTW1 = (X1 - 1) * (X2 - 1) + 1;
  static constexpr int NTW = L::TW2 ? NTWA + NTWB : NTW1;
  // kept psi(r) (KEEP variants): per state and plane, ROUNDS x X2 blocks of
  // KSLOT pair-row values, one per active pass-2 lane of the x round
  static constexpr int KSLOT = NTRX * X1;
  static constexpr int KPER = ROUNDS * X2 * KSLOT;
  static_assert(NX == NY && X1 == Y2 && X2 == Y1,
                "the fused plane passes need square x-y planes, y as R1 x R2, x as R2 x R1");
};

// P element of row y in a column (YA padding per Y1 rows)
template <int Y1, class L>
__device__ __forceinline__ int py(int y) {
  return L::YA ? y + L::YA * (y / Y1) : y;
}
// exchange element of position p in a row whose pass-1 blocks are R long
template <int R, class L>
__device__ __forceinline__ int pe(int p) {
  return L::EA ? p + L::EA * (p / R) : p;
}

// lane map: transform tr and thread j within it
template <bool TMIN, int G, int NTR>
__device__ __forceinline__ void lane_map(int tid, int& tr, int& j) {
  if (TMIN) {
    j = tid / NTR;
    tr = tid - j * NTR;
  } else {
    tr = tid / G;
    j = tid - tr * G;
  }
}

// pass-2 twiddle W_N^{t j}, conjugated for S = +1.  Table A: t < X1, j < X2
// (y passes, r2c pass 2); table B: t < X2, j < X1 (c2r pass 2).
template <typename T, int S, int RJ, class L>
__device__ __forceinline__ cpx<T> tw2(const cpx<T>* tw, int t, int j) {
  const cpx<T> w = L::TW2 ? tw[t * RJ + j] : tw[t * j];
  return {w.x, S < 0 ? w.y : -w.y};
}

// copy the twiddles and the per-column ky ranges into LDS (callers
// synchronise before use); the ranges are read for every state, and a
// global load there would cost a memory latency per state
template <typename T, int X1, int X2, int NT, class L>
__device__ __forceinline__ void load_twl(cpx<T>* tw, const double2* __restrict__ g, int2* ky,
                                         const FftCtx& c, int npx) {
  if (L::TW2) {
    for (int i = threadIdx.x; i < X1 * X2; i += NT) {
      const int t = i / X2, j = i - t * X2;  // A[t][j], t < X1
      tw[i] = {(T)g[t * j].x, (T)g[t * j].y};
      if (X1 != X2) {
        const int tb = i / X1, jb = i - tb * X1;  // B[t][j], t < X2
        tw[X1 * X2 + i] = {(T)g[tb * jb].x, (T)g[tb * jb].y};
      }
    }
  } else {
    for (int i = threadIdx.x; i < (X1 - 1) * (X2 - 1) + 1; i += NT)
      tw[i] = {(T)g[i].x, (T)g[i].y};
  }
  for (int i = threadIdx.x; i < npx; i += NT) ky[i] = make_int2(c.kylo[i], c.kyhi[i]);
}

// LDS bytes of either kernel: P (nplane_x columns) + E + twiddles + ky ranges
template <typename T, int X1, int X2, int Y1, int Y2, int NT, class L>
size_t plane3_smem_l(int npx) {
  using S = PS<X1, X2, Y1, Y2, NT, L>;
  return sizeof(cpx<T>) * ((size_t)npx * S::CY + S::EX + S::NTW) + sizeof(int2) * npx;
}

// y-FFT (FFTW_FORWARD) of the populated columns of one z-plane into P.
// zs: zbuf plane [kx][NY]; only ky in KY[kx] = [kylo, kyhi] of each column
// is read.  Pass 1 runs j-minor (loads along y), pass 2 in place.
template <typename T, int Y1, int Y2, int NT, class L>
__device__ __forceinline__ void y_fwd(const cpx<T>* __restrict__ zs, cpx<T>* P, const cpx<T>* TW,
                                      const int2* KY, int npx) {
  using S = PS<Y2, Y1, Y1, Y2, NT, L>;
  constexpr int NY = S::NY, GY = S::GY, NTRY = S::NTRY, CY = S::CY;
  const int tid = threadIdx.x;
  int tr, j, tr2, j2;
  lane_map<false, GY, NTRY>(tid, tr, j);
  lane_map<(L::MAPS & kMapYf2) != 0, GY, NTRY>(tid, tr2, j2);
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
        v[t] = (!Sph<NY, Y2>::cut(L::PR, t) && kyv >= lo && kyv <= hi) ? src[y]
                                                                          : cpx<T>{(T)0, (T)0};
      }
      Dft<T, Y1, -1>::run(v);
      cpx<T>* dst = P + kx * CY + j * (Y1 + L::YA);
#pragma unroll
      for (int t = 0; t < Y1; t++) dst[t] = v[t];
    }
  }
  __syncthreads();
  // pass 2 in place: thread j reads and writes the rows j + t*Y1 of its column
  for (int cb = 0; cb < npx; cb += NTRY) {
    const int kx = cb + tr2;
    if (tr2 < NTRY && kx < npx && j2 < Y1) {
      cpx<T>* col = P + kx * CY + j2;
      cpx<T> u[Y2];
#pragma unroll
      for (int t = 0; t < Y2; t++) u[t] = col[t * (Y1 + L::YA)];
#pragma unroll
      for (int t = 1; t < Y2; t++) u[t] = cm(u[t], tw2<T, -1, Y1, L>(TW, t, j2));
      Dft<T, Y2, -1>::run(u);
#pragma unroll
      for (int t = 0; t < Y2; t++) col[t * (Y1 + L::YA)] = u[t];
    }
  }
  __syncthreads();
}

// y-FFT (FFTW_BACKWARD) of the nplane_x columns of P into the zout plane,
// keeping ky in KY[kx]; E is the exchange buffer.  Pass 2 runs j-minor
// (stores along y).
template <typename T, int Y1, int Y2, int NT, class L>
__device__ __forceinline__ void y_bwd(const cpx<T>* P, cpx<T>* E, const cpx<T>* TW,
                                      cpx<T>* __restrict__ zo, const int2* KY, int npx) {
  using S = PS<Y2, Y1, Y1, Y2, NT, L>;
  constexpr int NY = S::NY, GY = S::GY, NTRY = S::NTRY, CY = S::CY, RY = S::RY;
  const int tid = threadIdx.x;
  int tr1, j1, tr, j;
  lane_map<(L::MAPS & kMapYb1) != 0, GY, NTRY>(tid, tr1, j1);
  lane_map<false, GY, NTRY>(tid, tr, j);
  for (int cb = 0; cb < npx; cb += NTRY) {
    {
      const int kx = cb + tr1;
      if (tr1 < NTRY && kx < npx && j1 < Y2) {
        cpx<T> v[Y1];
#pragma unroll
        for (int t = 0; t < Y1; t++) v[t] = P[kx * CY + py<Y1, L>(j1 + t * Y2)];
        Dft<T, Y1, +1>::run(v);
        cpx<T>* dst = E + tr1 * RY + j1 * (Y1 + L::EA);
#pragma unroll
        for (int t = 0; t < Y1; t++) dst[t] = v[t];
      }
    }
    __syncthreads();
    const int kx = cb + tr;
    if (tr < NTRY && kx < npx && j < Y1) {
      const cpx<T>* src = E + tr * RY + j;
      cpx<T> u[Y2];
#pragma unroll
      for (int t = 0; t < Y2; t++) u[t] = src[t * (Y1 + L::EA)];
#pragma unroll
      for (int t = 1; t < Y2; t++) u[t] = cm(u[t], tw2<T, +1, Y1, L>(TW, t, j));
      Dft<T, Y2, +1>::run(u);
      const int lo = KY[kx].x, hi = KY[kx].y;
      cpx<T>* dst = zo + (size_t)kx * NY;
#pragma unroll
      for (int t = 0; t < Y2; t++) {
        const int y = j + t * Y1;
        const int kyv = signed_k(y, NY);
        if (!Sph<NY, Y1>::cut(L::PR, t) && kyv >= lo && kyv <= hi) dst[y] = u[t];
      }
    }
    __syncthreads();
  }
}

// Lane maps of one x round: pass 1 of the c2r (thread j < X2 of pair-row q)
// and pass 2 with everything after it in registers (thread j < X1).
struct XLanes {
  int tr1, j1, q1;
  bool act1;
  int tr, j, q;
  bool act;
};

template <int X1, int X2, int Y1, int Y2, int NT, class L>
__device__ __forceinline__ XLanes x_lanes(int r) {
  using S = PS<X1, X2, Y1, Y2, NT, L>;
  XLanes x;
  lane_map<(L::MAPS & kMapXc1) != 0, S::GX, S::NTRX>(threadIdx.x, x.tr1, x.j1);
  lane_map<(L::MAPS & kMapXc2) != 0, S::GX, S::NTRX>(threadIdx.x, x.tr, x.j);
  x.q1 = r * S::NTRX + x.tr1;
  x.q = r * S::NTRX + x.tr;
  x.act1 = x.tr1 < S::NTRX && x.q1 < S::NPAIR;
  x.act = x.tr < S::NTRX && x.q < S::NPAIR;
  return x;
}

// Slot of an active pass-2 lane (x.tr < NTRX, x.j < X1) in a kept psi(r)
// block: the lane's own thread index under the transform-minor map, so a
// wave's loads and stores of one t run along consecutive elements.  The
// slot does not depend on the round.
template <int X1, int NTRX, class L>
__device__ __forceinline__ int kslot(const XLanes& x) {
  return (L::MAPS & kMapXc2) ? x.j * NTRX + x.tr : x.tr * X1 + x.j;
}

// One round of Hermitian x-FFTs (c2r by the pair-row trick) from the rows of
// P: pair q carries rows y1 = 2q (real part) and y2 = 2q+1 (imaginary part).
// On return thread (x.tr, x.j < X1) holds psi(r) at x = j + t*X1 in u[t].
// Ends after the pass-2 reads of E (callers sync before reusing E).
template <typename T, int X1, int X2, int Y1, int Y2, int NT, class L>
__device__ __forceinline__ void x_c2r(const cpx<T>* P, cpx<T>* E, const cpx<T>* TW, int npx,
                                      const XLanes& x, cpx<T> (&u)[X2]) {
  using S = PS<X1, X2, Y1, Y2, NT, L>;
  constexpr int NX = S::NX, NY = S::NY, CY = S::CY, RE = S::RE;
  if (x.act1 && x.j1 < X2) {
    const int y1 = 2 * x.q1, y2 = y1 + 1;
    const int o1 = py<Y1, L>(y1), o2 = py<Y1, L>(y2);
    cpx<T> v[X1];
#pragma unroll
    for (int t = 0; t < X1; t++) {
      const int k = x.j1 + t * X2;
      cpx<T> Z = {(T)0, (T)0};
      int kk = -1;
      bool mirror = false;
      if (Sph<NX, X2>::cut(L::PR, t)) {
      } else if (k < npx) {
        kk = k;
      } else if (k > NX - npx) {
        kk = NX - k;
        mirror = true;
      }
      if (kk >= 0) {
        const cpx<T> A = P[kk * CY + o1];
        const cpx<T> B = (y2 < NY) ? P[kk * CY + o2] : cpx<T>{(T)0, (T)0};
        if (kk == 0) Z = {A.x, B.x};
        else if (!mirror) Z = {A.x - B.y, A.y + B.x};
        else Z = {A.x + B.y, B.x - A.y};
      }
      v[t] = Z;
    }
    Dft<T, X1, -1>::run(v);
    cpx<T>* dst = E + x.tr1 * RE + x.j1 * (X1 + L::EA);
#pragma unroll
    for (int t = 0; t < X1; t++) dst[t] = v[t];
  }
  __syncthreads();
  if (x.act && x.j < X1) {
    const cpx<T>* src = E + x.tr * RE + x.j;
#pragma unroll
    for (int t = 0; t < X2; t++) u[t] = src[t * (X1 + L::EA)];
#pragma unroll
    for (int t = 1; t < X2; t++) u[t] = cm(u[t], tw2<T, -1, X1, L>(TW + S::NTWA * (X1 != X2 && L::TW2), t, x.j));
    Dft<T, X2, -1>::run(u);
  }
}

// G->R plane pass.  Grid (nz, row parts, state chunks).  Row part rp covers
// x rounds [rp*RPW, rp*RPW + RPW).  PSIR: write psi(r) of state
// c.psir_state into c.psir ([Nz][Ny][Nx] of T) instead of rho partials.
// KEEP: also write every state's psi(r) to c.psk ([nst][nz][KPER] complex T,
// pair rows as real and imaginary parts) for the R->G pass to read back
// (the real-space plane the reference's RealSpacePlane chare holds between
// the density and V psi).
template <typename T, int X1, int X2, int Y1, int Y2, int NT, int RPW, bool PSIR, class L,
          bool KEEP = false>
__global__ __launch_bounds__(NT) void k3_g2r(FftCtx c) {
  using S = PS<X1, X2, Y1, Y2, NT, L>;
  constexpr int NX = S::NX, NY = S::NY;
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
  const size_t plane = (size_t)npx * NY;
  const size_t NP = (size_t)NX * NY;
  load_twl<T, X1, X2, NT, L>(TW, c.twx, KY, c, npx);
  __syncthreads();
  double acc[RPW][2 * X2];
#pragma unroll
  for (int a = 0; a < RPW; a++)
#pragma unroll
    for (int b = 0; b < 2 * X2; b++) acc[a][b] = 0.0;
  for (int s = st0; s < st1; s++) {
    const cpx<T>* zs = (const cpx<T>*)c.zbuf + ((size_t)s * c.g.nz + z) * plane;
    y_fwd<T, Y1, Y2, NT, L>(zs, P, TW, KY, npx);
    const double oc = c.occ[s];
#pragma unroll
    for (int rr = 0; rr < RPW; rr++) {
      const int r = r0 + rr;
      if (r < S::ROUNDS) {
        const XLanes x = x_lanes<X1, X2, Y1, Y2, NT, L>(r);
        cpx<T> u[X2];
        x_c2r<T, X1, X2, Y1, Y2, NT, L>(P, E, TW, npx, x, u);
        if (x.act && x.j < X1) {
          const int y1 = 2 * x.q, y2 = y1 + 1;
          if (KEEP) {
            cpx<T>* kp = (cpx<T>*)c.psk +
                         (((size_t)s * c.g.nz + z) * S::ROUNDS + r) * (X2 * S::KSLOT) +
                         kslot<X1, S::NTRX, L>(x);
#pragma unroll
            for (int t = 0; t < X2; t++) kp[t * S::KSLOT] = u[t];
          }
          if (PSIR) {
            T* pr = (T*)c.psir + (size_t)z * NP;
#pragma unroll
            for (int t = 0; t < X2; t++) {
              pr[(size_t)y1 * NX + x.j + t * X1] = u[t].x;
              if (y2 < NY) pr[(size_t)y2 * NX + x.j + t * X1] = u[t].y;
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
    const XLanes x = x_lanes<X1, X2, Y1, Y2, NT, L>(r);
    if (r < S::ROUNDS && x.act && x.j < X1) {
      const int y1 = 2 * x.q, y2 = y1 + 1;
#pragma unroll
      for (int t = 0; t < X2; t++) {
        rp_out[(size_t)y1 * NX + x.j + t * X1] = acc[rr][2 * t];
        if (y2 < NY) rp_out[(size_t)y2 * NX + x.j + t * X1] = acc[rr][2 * t + 1];
      }
    }
  }
}

// R->G plane pass.  Grid (nz, state chunks).  VREG: V(r) of the thread's
// pass-2 positions for all rounds is held in registers across the chunk.
// FP32 up to N = 100: at most 168 VGPRs (154 at N = 100, 144 pruned), three
// waves per SIMD (measured faster than two; larger planes would spill).  RHO: also accumulate rho
// partials of the chunk from the recomputed psi(r), for a V(r) that does
// not depend on this step's rho (then no separate G->R pass is needed).
// KEEP: psi(r) is read back from c.psk, as the G->R pass with KEEP wrote it,
// instead of being recomputed from zbuf (no y-FFT and no x c2r here); each
// round's values are loaded one round ahead.
template <typename T, int X1, int X2, int Y1, int Y2, int NT, bool VREG, bool RHO, class L,
          bool KEEP = false>
__global__ __launch_bounds__(NT, (sizeof(T) == 4 && !RHO) ? (X1 * X2 <= 100 ? 3 : 2) : 1) void
k3_r2g(FftCtx c) {
  using S = PS<X1, X2, Y1, Y2, NT, L>;
  constexpr int NX = S::NX, NY = S::NY, NTRX = S::NTRX, R = S::ROUNDS;
  constexpr int CY = S::CY, RE = S::RE;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int npx = c.g.nplane_x;
  cpx<T>* P = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* E = P + (size_t)npx * CY;
  cpx<T>* TW = E + S::EX;
  int2* KY = reinterpret_cast<int2*>(TW + S::NTW);
  const int z = blockIdx.x, chunk = blockIdx.y;
  const int st0 = chunk * c.cs_r, st1 = min(c.g.nst, st0 + c.cs_r);
  const int tid = threadIdx.x;
  const size_t plane = (size_t)npx * NY;
  const size_t NP = (size_t)NX * NY;
  const double* vz = c.vks + (size_t)z * NP;
  load_twl<T, X1, X2, NT, L>(TW, c.twx, KY, c, npx);
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
      const XLanes x = x_lanes<X1, X2, Y1, Y2, NT, L>(r);
      const bool act = x.act && x.j < X1;
      const int ya = 2 * min(x.q, S::NPAIR - 1), yb = min(ya + 1, NY - 1);
#pragma unroll
      for (int t = 0; t < X2; t++) {
        const int xx = min(x.j + t * X1, NX - 1);
        const T a = (T)vz[(size_t)ya * NX + xx], b = (T)vz[(size_t)yb * NX + xx];
        vr[r][2 * t] = act ? a : (T)0;
        vr[r][2 * t + 1] = (act && 2 * x.q + 1 < NY) ? b : (T)0;
      }
    }
  }
  int trf, jf;  // r2c pass-2 lanes
  lane_map<(L::MAPS & kMapRx2) != 0, S::GX, NTRX>(tid, trf, jf);
  // split loop: element idx = qq npx + k of a round, stepped by NT without a
  // division per element (idx + NT = (qq + sdq) npx + k + sdk, k + sdk < 2 npx)
  const int sq0 = tid / npx, sk0 = tid - sq0 * npx;
  const int sdq = NT / npx, sdk = NT - sdq * npx;
  // KEEP: this thread's slot in the kept blocks (slot 0 for lanes without
  // pass-2 work, whose values are never used), and the next round's values
  const cpx<T>* kbase = nullptr;
  cpx<T> uk[KEEP ? X2 : 1];
  if (KEEP) {
    const XLanes x0 = x_lanes<X1, X2, Y1, Y2, NT, L>(0);
    const int slot = (x0.tr < NTRX && x0.j < X1) ? kslot<X1, NTRX, L>(x0) : 0;
    kbase = (const cpx<T>*)c.psk + (size_t)z * S::KPER + slot;
    if (st0 < st1) {
      const cpx<T>* kp = kbase + (size_t)st0 * c.g.nz * S::KPER;
#pragma unroll
      for (int t = 0; t < X2; t++) uk[KEEP ? t : 0] = kp[t * S::KSLOT];
    }
  }
  for (int s = st0; s < st1; s++) {
    if (!KEEP) {
      const cpx<T>* zs = (const cpx<T>*)c.zbuf + ((size_t)s * c.g.nz + z) * plane;
      y_fwd<T, Y1, Y2, NT, L>(zs, P, TW, KY, npx);
    }
    const T oc = (T)c.occ[s];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const XLanes x = x_lanes<X1, X2, Y1, Y2, NT, L>(r);
      // without VREG: this round's V, loaded (branch-free) ahead of the transform
      T vv[VREG ? 1 : 2 * X2];
      if (!VREG) {
        const int ya = 2 * min(x.q, S::NPAIR - 1), yb = min(ya + 1, NY - 1);
#pragma unroll
        for (int t = 0; t < X2; t++) {
          const int xx = min(x.j + t * X1, NX - 1);
          vv[VREG ? 0 : 2 * t] = (T)vz[(size_t)ya * NX + xx];
          vv[VREG ? 0 : 2 * t + 1] = (T)vz[(size_t)yb * NX + xx];
        }
      }
      cpx<T> u[X2];
      if (KEEP) {
#pragma unroll
        for (int t = 0; t < X2; t++) u[t] = uk[KEEP ? t : 0];
        // the following round, or round 0 of the next state (uniform branch)
        const int rn = r + 1 < R ? r + 1 : 0, sn = r + 1 < R ? s : s + 1;
        if (sn < st1) {
          const cpx<T>* kp =
              kbase + ((size_t)sn * c.g.nz * S::ROUNDS + rn) * (X2 * S::KSLOT);
#pragma unroll
          for (int t = 0; t < X2; t++) uk[KEEP ? t : 0] = kp[t * S::KSLOT];
        }
      } else {
        x_c2r<T, X1, X2, Y1, Y2, NT, L>(P, E, TW, npx, x, u);
        __syncthreads();  // E is rewritten below
      }
      if (x.act && x.j < X1) {
        const int y2 = 2 * x.q + 1;
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
        cpx<T>* dst = E + x.tr * RE + x.j * (X2 + L::EA);
#pragma unroll
        for (int t = 0; t < X2; t++) dst[t] = w[t];
      }
      __syncthreads();
      const int qf = r * NTRX + trf;
      if (trf < NTRX && qf < S::NPAIR && jf < X2) {
        cpx<T>* row = E + trf * RE + jf;
        cpx<T> a[X1];
#pragma unroll
        for (int t = 0; t < X1; t++) a[t] = row[t * (X2 + L::EA)];
#pragma unroll
        for (int t = 1; t < X1; t++) a[t] = cm(a[t], tw2<T, -1, X2, L>(TW, t, jf));
        Dft<T, X1, -1>::run(a);
#pragma unroll
        for (int t = 0; t < X1; t++)
          if (!Sph<NX, X2>::cut(L::PR, t)) row[t * (X2 + L::EA)] = a[t];
      }
      __syncthreads();
      // split the two real rows, keep kx < nplane_x, conjugated, into P
      const int nq = min(NTRX, S::NPAIR - r * NTRX);
      for (int idx = tid, qq = sq0, k = sk0; idx < nq * npx;
           idx += NT, qq += sdq + (k + sdk >= npx), k += sdk - (k + sdk >= npx ? npx : 0)) {
        const int r1 = 2 * (r * NTRX + qq), r2 = r1 + 1;
        const cpx<T> Zk = E[qq * RE + pe<X2, L>(k)];
        const cpx<T> Zm = E[qq * RE + pe<X2, L>(k == 0 ? 0 : NX - k)];
        P[k * CY + py<Y1, L>(r1)] = {(T)0.5 * (Zk.x + Zm.x), -(T)0.5 * (Zk.y - Zm.y)};
        if (r2 < NY) P[k * CY + py<Y1, L>(r2)] = {(T)0.5 * (Zk.y + Zm.y), (T)0.5 * (Zk.x - Zm.x)};
      }
      __syncthreads();
    }
    cpx<T>* zo = (cpx<T>*)c.zout + ((size_t)s * c.g.nz + z) * plane;
    y_bwd<T, Y1, Y2, NT, L>(P, E, TW, zo, KY, npx);
  }
  if (!RHO) return;
  // rho partial of this chunk (same layout and per-element order as k3_g2r)
  double* rp_out = c.rho_part + (size_t)chunk * NP * c.g.nz + (size_t)z * NP;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const XLanes x = x_lanes<X1, X2, Y1, Y2, NT, L>(r);
    if (x.act && x.j < X1) {
      const int y1 = 2 * x.q, y2 = y1 + 1;
#pragma unroll
      for (int t = 0; t < X2; t++) {
        rp_out[(size_t)y1 * NX + x.j + t * X1] = acc[r][2 * t];
        if (y2 < NY) rp_out[(size_t)y2 * NX + x.j + t * X1] = acc[r][2 * t + 1];
      }
    }
  }
}

// rho accumulation fits the R->G pass's registers: at most two rounds of
// radix-10 pair rows
template <int X1, int X2, int NT>
constexpr bool r2g_rho_ok() {
  return (((X1 * X2 + 1) / 2 + NT / (X1 > X2 ? X1 : X2) - 1) / (NT / (X1 > X2 ? X1 : X2))) * X2 <=
         20;
}

// ------------------------------------------------------------------ launch

enum PlaneKind { G2R = 0, R2G = 1, G2R_PSIR = 2, R2G_RHO = 3 };

constexpr int kPlaneThreads = 256;

// Layouts.  OldLay is the layout before the bank-conflict work (one
// twiddle table by t*j, j-minor lanes everywhere, P columns NY|1, unpadded
// exchange rows); PadLay has the [t][j] twiddle tables, transform-minor
// lanes wherever only LDS is touched and odd P column and exchange row
// strides.  Measured on the 100^3 / 128-state R->G pass (FP32): 286 -> 256
// us, identical results.  Padding grows the LDS a little, so a launch uses
// PadLay only where that keeps the workgroups per CU of OldLay (FP64 at
// 100^2 planes sits 8 bytes under half a CU's LDS with OldLay).
template <int N, bool PR = false>
using OldLay = Lay<(N % 2 == 0) ? 1 : 0, 0, 0, 0, 0, false, PR>;
template <int N, bool PR = false>
using PadLay = Lay<(N % 2 == 0) ? 5 : 4, 0, 0, (N % 2 == 0) ? 5 : 4, 31, true, PR>;

// x rounds per k3_g2r workgroup: the rho accumulators are 2*X2 doubles per
// round and thread; FP32 keeps two rounds, FP64 two when pruned to the
// sphere (220 VGPRs at N = 100, no spills) and one otherwise, which makes
// two row parts at N = 100 that each read zbuf and repeat the y-FFT of the
// plane.  When the LDS of one
// workgroup exceeds half a CU's (one workgroup per CU, four waves), the
// register file has room for every round and one workgroup covers the whole
// plane ("wide"), so the y-FFT of a plane is not repeated per row part.
template <typename T, class L>
constexpr int row_rounds() {
  return sizeof(T) == 8 && !L::PR ? 1 : 2;
}

constexpr size_t kCuLds = 160 * 1024;

template <typename T, int X1, int X2, int Y1, int Y2, int NT, class L>
bool g2r_wide(int npx) {
  return 2 * plane3_smem_l<T, X1, X2, Y1, Y2, NT, L>(npx) > kCuLds;
}

// PadLay where it keeps OldLay's workgroups per CU (by LDS)
template <typename T, int X1, int X2, int Y1, int Y2, int NT>
bool use_pad(int npx) {
  const size_t so = plane3_smem_l<T, X1, X2, Y1, Y2, NT, OldLay<X1 * X2>>(npx);
  const size_t sp = plane3_smem_l<T, X1, X2, Y1, Y2, NT, PadLay<X1 * X2>>(npx);
  return kCuLds / sp >= kCuLds / so;
}

// Wide workgroups: FP64 planes of 108..128 points use more than half a CU's
// LDS with 256 threads, so one workgroup (four waves) runs per CU.  With 512
// threads the plane's state and exchange buffers still fit one CU, twice as
// many transforms run per round, and the CU holds eight waves.
constexpr int kWideThreads = 512;
template <typename T, int X1, int X2, int Y1, int Y2>
bool use_wide(int npx) {
  if constexpr (sizeof(T) == 8 && X1 * X2 >= 96) {
    using L = OldLay<X1 * X2>;
    return 2 * plane3_smem_l<T, X1, X2, Y1, Y2, kPlaneThreads, L>(npx) > kCuLds &&
           plane3_smem_l<T, X1, X2, Y1, Y2, kWideThreads, L>(npx) <= kCuLds;
  }
  return false;
}

template <typename T, int X1, int X2, int Y1, int Y2, int NT>
size_t plane3_smem(int npx) {
  if (use_wide<T, X1, X2, Y1, Y2>(npx))
    return plane3_smem_l<T, X1, X2, Y1, Y2, kWideThreads, OldLay<X1 * X2>>(npx);
  return use_pad<T, X1, X2, Y1, Y2, NT>(npx)
             ? plane3_smem_l<T, X1, X2, Y1, Y2, NT, PadLay<X1 * X2>>(npx)
             : plane3_smem_l<T, X1, X2, Y1, Y2, NT, OldLay<X1 * X2>>(npx);
}

// Launch facts of one size for the host's chunk choice: row parts of the
// G->R grid and resident workgroups per CU of both kernels.
template <typename T, int X1, int X2, int Y1, int Y2, int NT, class L>
hipError_t plane3_info_l(int npx, int* rowparts, int* occ_g2r, int* occ_r2g, int* rho_ok) {
  using S = PS<X1, X2, Y1, Y2, NT, L>;
  constexpr bool VREG = sizeof(T) == 4 && S::ROUNDS * 2 * X2 <= 48;
  const size_t smem = plane3_smem_l<T, X1, X2, Y1, Y2, NT, L>(npx);
  const bool wide = g2r_wide<T, X1, X2, Y1, Y2, NT, L>(npx);
  *rowparts = wide ? 1 : (S::ROUNDS + row_rounds<T, L>() - 1) / row_rounds<T, L>();
  const void* g = wide ? (const void*)k3_g2r<T, X1, X2, Y1, Y2, NT, S::ROUNDS, false, L>
                       : (const void*)k3_g2r<T, X1, X2, Y1, Y2, NT, row_rounds<T, L>(), false, L>;
  const void* r = (const void*)k3_r2g<T, X1, X2, Y1, Y2, NT, VREG, false, L>;
  // the R->G pass with rho accumulators spills at 512 threads
  *rho_ok = r2g_rho_ok<X1, X2, NT>() && NT == kPlaneThreads ? 1 : 0;
  hipError_t e;
  if ((e = oa::grant_lds(g, smem)) != hipSuccess || (e = oa::grant_lds(r, smem)) != hipSuccess)
    return e;
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(occ_g2r, g, NT, smem)) != hipSuccess)
    return e;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ_r2g, r, NT, smem);
}

template <typename T, int X1, int X2, int Y1, int Y2, int NT, class L>
hipError_t launch_plane3_l(int kind, const FftCtx& c, hipStream_t st) {
  using S = PS<X1, X2, Y1, Y2, NT, L>;
  constexpr int kRowRounds = row_rounds<T, L>();
  constexpr int RPn = (S::ROUNDS + kRowRounds - 1) / kRowRounds;
  constexpr bool VREG = sizeof(T) == 4 && S::ROUNDS * 2 * X2 <= 48;
  const size_t smem = plane3_smem_l<T, X1, X2, Y1, Y2, NT, L>(c.g.nplane_x);
  if (c.g.nx != S::NX || c.g.ny != S::NY) return hipErrorInvalidValue;
  // raise the dynamic LDS limit of an instantiation (once per device)
  auto grant = [&](const void* fn) -> hipError_t { return oa::grant_lds(fn, smem); };
  hipError_t e = hipSuccess;
  switch (kind) {
    case G2R: {
      // c.psk set: the KEEP variant also writes psi(r) for the R->G pass
      const bool wide = g2r_wide<T, X1, X2, Y1, Y2, NT, L>(c.g.nplane_x);
      auto fn = wide ? (c.psk ? k3_g2r<T, X1, X2, Y1, Y2, NT, S::ROUNDS, false, L, true>
                              : k3_g2r<T, X1, X2, Y1, Y2, NT, S::ROUNDS, false, L>)
                     : (c.psk ? k3_g2r<T, X1, X2, Y1, Y2, NT, kRowRounds, false, L, true>
                              : k3_g2r<T, X1, X2, Y1, Y2, NT, kRowRounds, false, L>);
      if ((e = grant((const void*)fn)) != hipSuccess) return e;
      hipLaunchKernelGGL(fn, dim3(c.g.nz, wide ? 1 : RPn, c.nchunk), dim3(NT), smem, st, c);
      break;
    }
    case G2R_PSIR: {
      auto fn = k3_g2r<T, X1, X2, Y1, Y2, NT, kRowRounds, true, L>;
      if ((e = grant((const void*)fn)) != hipSuccess) return e;
      hipLaunchKernelGGL(fn, dim3(c.g.nz, RPn, 1), dim3(NT), smem, st, c);
      break;
    }
    case R2G: {
      // c.psk set: psi(r) read back from the G->R pass's KEEP output
      auto fn = c.psk ? k3_r2g<T, X1, X2, Y1, Y2, NT, VREG, false, L, true>
                      : k3_r2g<T, X1, X2, Y1, Y2, NT, VREG, false, L>;
      if ((e = grant((const void*)fn)) != hipSuccess) return e;
      hipLaunchKernelGGL(fn, dim3(c.g.nz, c.nchunk_r), dim3(NT), smem, st, c);
      break;
    }
    case R2G_RHO: {
      constexpr bool kRho = r2g_rho_ok<X1, X2, NT>() && NT == kPlaneThreads;
      if (!kRho) return hipErrorInvalidValue;
      auto fn = k3_r2g<T, X1, X2, Y1, Y2, NT, VREG, kRho, L>;
      if ((e = grant((const void*)fn)) != hipSuccess) return e;
      hipLaunchKernelGGL(fn, dim3(c.g.nz, c.nchunk_r), dim3(NT), smem, st, c);
      break;
    }
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// complex elements per state and plane of the kept psi(r) (c.psk) of the
// variant a launch takes (the block layout depends only on the threads)
template <typename T, int X1, int X2, int Y1, int Y2>
size_t plane3_kper(int npx) {
  constexpr int N = X1 * X2;
  if (use_wide<T, X1, X2, Y1, Y2>(npx))
    return PS<X1, X2, Y1, Y2, kWideThreads, OldLay<N>>::KPER;
  return PS<X1, X2, Y1, Y2, kPlaneThreads, OldLay<N>>::KPER;
}

// the variant a launch takes: 512 threads (OldLay) for wide FP64 planes,
// else 256 with PadLay where it keeps OldLay's workgroups per CU; PR: pruned
template <typename T, int X1, int X2, int Y1, int Y2, bool PR>
hipError_t plane3_info_pr(int npx, int* rowparts, int* occ_g2r, int* occ_r2g, int* rho_ok) {
  constexpr int N = X1 * X2, NT = kPlaneThreads;
  if constexpr (sizeof(T) == 8 && N >= 96) {
    if (use_wide<T, X1, X2, Y1, Y2>(npx))
      return plane3_info_l<T, X1, X2, Y1, Y2, kWideThreads, OldLay<N, PR>>(npx, rowparts, occ_g2r,
                                                                            occ_r2g, rho_ok);
  }
  return use_pad<T, X1, X2, Y1, Y2, NT>(npx)
             ? plane3_info_l<T, X1, X2, Y1, Y2, NT, PadLay<N, PR>>(npx, rowparts, occ_g2r, occ_r2g,
                                                                   rho_ok)
             : plane3_info_l<T, X1, X2, Y1, Y2, NT, OldLay<N, PR>>(npx, rowparts, occ_g2r, occ_r2g,
                                                                   rho_ok);
}

template <typename T, int X1, int X2, int Y1, int Y2>
hipError_t plane3_info(int npx, bool prune, int* rowparts, int* occ_g2r, int* occ_r2g,
                       int* rho_ok) {
  return prune ? plane3_info_pr<T, X1, X2, Y1, Y2, true>(npx, rowparts, occ_g2r, occ_r2g, rho_ok)
               : plane3_info_pr<T, X1, X2, Y1, Y2, false>(npx, rowparts, occ_g2r, occ_r2g, rho_ok);
}

template <typename T, int X1, int X2, int Y1, int Y2, bool PR>
hipError_t launch_plane3_pr(int kind, const FftCtx& c, hipStream_t st) {
  constexpr int N = X1 * X2, NT = kPlaneThreads;
  const int npx = c.g.nplane_x;
  if constexpr (sizeof(T) == 8 && N >= 96) {
    if (use_wide<T, X1, X2, Y1, Y2>(npx))
      return launch_plane3_l<T, X1, X2, Y1, Y2, kWideThreads, OldLay<N, PR>>(kind, c, st);
  }
  return use_pad<T, X1, X2, Y1, Y2, NT>(npx)
             ? launch_plane3_l<T, X1, X2, Y1, Y2, NT, PadLay<N, PR>>(kind, c, st)
             : launch_plane3_l<T, X1, X2, Y1, Y2, NT, OldLay<N, PR>>(kind, c, st);
}

// c.prune (set by the host when every populated |kx|, |ky| <= (N - 1) / 4)
// selects the variants pruned to the sphere
template <typename T, int X1, int X2, int Y1, int Y2>
hipError_t launch_plane3(int kind, const FftCtx& c, hipStream_t st) {
  return c.prune ? launch_plane3_pr<T, X1, X2, Y1, Y2, true>(kind, c, st)
                 : launch_plane3_pr<T, X1, X2, Y1, Y2, false>(kind, c, st);
}

}  // namespace oa3
