// oa_gemm.hip -- dense FP64 contractions of the CP electronic step on
// CDNA4 matrix cores (v_mfma_f64_16x16x4f64).
//
//   gram    : S = 2 Re Psi^H Psi over stored points with kx=0 halved
//             (PairCalculator::multiplyForward, ckPairCalculator.C:682-802)
//             and Lambda = Re Psi^H F (asymmetric PairCalculator, alpha = 1).
//             One workgroup owns a 128 x 128 output block and one slice of
//             the long point axis, so each slice of Psi is read once per
//             block; the slices' partial blocks are summed in a fixed order.
//   rotate  : Psi' = Psi T and F -= w Psi Lambda
//             (bwMultiplyHelper, ckPairCalculator.C:1189-1425;
//              unpackLambda, CP_State_GSpacePlane.C:2045-2225), optionally
//             followed in the same pass by the steepest-descent update
//             psi += dt/m F (CP_integrate_min_STD, cp_min_std.C:56-86).
//   ns      : Newton-Schulz iteration for (S/2)^{-1/2}
//             (Ortho::do_iteration/step_2/step_3/tolerance_check,
//              ortho.C:862-985; serial twin cp_lowdin.C:180-265): one
//             cooperative launch that loops on the device with grid
//             barriers, operand panels staged in LDS; fallback: two launches
//             per iteration with a device-side early exit.
//
// MFMA operand layout (measured on gfx950): lane l supplies A[l%16][l/16]
// and B[l/16][l%16] and receives D[l/16 + 4v][l%16] in element v.  Within
// one MFMA the assignment of the four k values to the four lane groups is
// free as long as A and B agree: in the gram, lane group g takes the
// contiguous run k0+4g..k0+4g+3 over four consecutive MFMAs, so every lane
// issues one 32-byte load per operand row and 16 k values.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <type_traits>

#include "oa_common.h"
#include "oa_gemm_common.h"
#include "oa_kernels.h"

namespace oa {

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma16(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v4d vzero() { return (v4d){0.0, 0.0, 0.0, 0.0}; }

// ================================================================== gram
constexpr int kGB = 128;        // output block edge

// Diagonal block of the symmetric product: its 36 upper tiles (ti <= tj) in
// four groups with few distinct row fragments per wave:
//   wave 0: ti <= tj < 4           (10 tiles, row tiles 0-3)
//   wave 1: ti in {0,1}, tj >= 4   ( 8 tiles, row tiles 0,1,4-7)
//   wave 2: ti in {2,3}, tj >= 4   ( 8 tiles, row tiles 2,3,4-7)
//   wave 3: 4 <= ti <= tj          (10 tiles, row tiles 4-7)
// Fragment slot f of wave W holds row tile dg_rt(W, f); tile t of wave W is
// the product of slots dg_a(W, t) and dg_b(W, t).
__host__ __device__ constexpr int up_i(int idx, int m) {
  int i = 0;
  while (idx >= m - i) {
    idx -= m - i;
    i++;
  }
  return i;
}
__host__ __device__ constexpr int up_j(int idx, int m) {
  int i = 0;
  while (idx >= m - i) {
    idx -= m - i;
    i++;
  }
  return i + idx;
}
__host__ __device__ constexpr int dg_nf(int W) { return (W == 0 || W == 3) ? 4 : 6; }
__host__ __device__ constexpr int dg_nt(int W) { return (W == 0 || W == 3) ? 10 : 8; }
__host__ __device__ constexpr int dg_rt(int W, int f) {
  return W == 0 ? f : W == 3 ? 4 + f : W == 2 ? 2 + f : (f < 2 ? f : 2 + f);
}
__host__ __device__ constexpr int dg_a(int W, int t) {
  return (W == 0 || W == 3) ? up_i(t, 4) : t / 4;
}
__host__ __device__ constexpr int dg_b(int W, int t) {
  return (W == 0 || W == 3) ? up_j(t, 4) : 2 + t % 4;
}

// f[e] = X[row][kk + e]: one 32-byte load, kk a multiple of 4 inside the
// row's allocation (chunks start on multiples of 16 below K <= lda, lda a
// multiple of 16).  Unconditional: rows past the matrix are clamped by the
// caller (their outputs are never stored) and elements outside the slice's
// region are zeroed where they are consumed (gmask), so no load sits in a
// branch and the ring's loads stay in flight.
__device__ __forceinline__ void gfrag(const double* __restrict__ X, int lda, int row, int kk,
                                      double (&f)[4]) {
  const double4 v = *reinterpret_cast<const double4*>(X + (size_t)row * lda + kk);
  f[0] = v.x;
  f[1] = v.y;
  f[2] = v.z;
  f[3] = v.w;
}

// zero f[e] for kk + e outside [rs, re)
__device__ __forceinline__ void gmask(int kk, int rs, int re, double (&f)[4]) {
#pragma unroll
  for (int e = 0; e < 4; e++) f[e] = (kk + e >= rs && kk + e < re) ? f[e] : 0.0;
}

template <int NF>
struct Frag {
  double f[NF][4];
};
constexpr int kGramRing = 3;

// Diagonal block of the symmetric product, wave W (tile groups above); the
// row fragments serve as both operands.
template <int W>
__device__ __forceinline__ void gram_diag_wave(const GramArgs& a, const GSlice& s, int sp,
                                               int bi) {
  constexpr int NF = dg_nf(W), NTW = dg_nt(W);
  const int lane = threadIdx.x & 63, l16 = lane & 15, g = lane >> 4;
  int row[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) {
    const int r = bi * kGB + 16 * dg_rt(W, f) + l16;
    row[f] = r < a.n ? r : 0;  // rows past n: outputs not stored
  }
  v4d acc[NTW];
#pragma unroll
  for (int t = 0; t < NTW; t++) acc[t] = vzero();
  auto mfmas = [&](const double (&f)[NF][4]) __attribute__((always_inline)) {
#pragma unroll
    for (int e = 0; e < 4; e++)
      static_for<0, NTW>([&](auto tc) __attribute__((always_inline)) {
        constexpr int t = decltype(tc)::value;
        constexpr int ia = dg_a(W, t), ib = dg_b(W, t);
        acc[t] = mfma16(f[ia][e], f[ib][e], acc[t]);
      });
  };
  Frag<NF> ring[kGramRing];
  chunk_ring(
      s.c0, s.c1, ring,
      [&](int c, Frag<NF>& d) __attribute__((always_inline)) {
        const int kk = s.base + 16 * c + 4 * g;
#pragma unroll
        for (int f = 0; f < NF; f++) gfrag(a.A, a.lda, row[f], kk, d.f[f]);
      },
      [&](int c, const Frag<NF>& d) __attribute__((always_inline)) {
        const int k0 = s.base + 16 * c;
        if (k0 >= s.rs && k0 + 16 <= s.re) {
          mfmas(d.f);
        } else {  // a chunk at the region's edge (both operands are these rows)
          Frag<NF> m = d;
#pragma unroll
          for (int f = 0; f < NF; f++) gmask(k0 + 4 * g, s.rs, s.re, m.f[f]);
          mfmas(m.f);
        }
      });
  double* part = a.part + (size_t)sp * a.n * a.n;
#pragma unroll
  for (int t = 0; t < NTW; t++) {
    const int r0 = bi * kGB + 16 * dg_rt(W, dg_a(W, t));
    const int c0 = bi * kGB + 16 * dg_rt(W, dg_b(W, t));
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = r0 + g + 4 * v, col = c0 + l16;
      if (r < a.n && col < a.n) part[(size_t)r * a.n + col] = s.w * acc[t][v];
    }
  }
}

// Full block: wave (wr, wc) owns the 64 x 64 quadrant (4 x 4 tiles).
__device__ __forceinline__ void gram_full_wave(const GramArgs& a, const double* __restrict__ B,
                                               const GSlice& s, int sp, int bi, int bj) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, l16 = lane & 15, g = lane >> 4;
  const int wr = wave >> 1, wc = wave & 1;
  const int rbase = bi * kGB + wr * 64, cbase = bj * kGB + wc * 64;
  if (rbase >= a.n || cbase >= a.nb) return;
  int ra[4], rb[4];
#pragma unroll
  for (int x = 0; x < 4; x++) {  // rows past the matrices: outputs not stored
    const int r = rbase + 16 * x + l16, q = cbase + 16 * x + l16;
    ra[x] = r < a.n ? r : 0;
    rb[x] = q < a.nb ? q : 0;
  }
  v4d acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) acc[x][y] = vzero();
  auto mfmas = [&](const double (&f)[8][4]) __attribute__((always_inline)) {
#pragma unroll
    for (int e = 0; e < 4; e++)
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) acc[x][y] = mfma16(f[x][e], f[4 + y][e], acc[x][y]);
  };
  Frag<8> ring[kGramRing];  // fragments 0-3: rows of A, 4-7: rows of B
  chunk_ring(
      s.c0, s.c1, ring,
      [&](int c, Frag<8>& d) __attribute__((always_inline)) {
        const int kk = s.base + 16 * c + 4 * g;
#pragma unroll
        for (int x = 0; x < 4; x++) {
          gfrag(a.A, a.lda, ra[x], kk, d.f[x]);
          gfrag(B, a.lda, rb[x], kk, d.f[4 + x]);
        }
      },
      [&](int c, const Frag<8>& d) __attribute__((always_inline)) {
        const int k0 = s.base + 16 * c;
        if (k0 >= s.rs && k0 + 16 <= s.re) {
          mfmas(d.f);
        } else {  // a chunk at the region's edge (padding past K may hold anything)
          Frag<8> m = d;
#pragma unroll
          for (int x = 0; x < 8; x++) gmask(k0 + 4 * g, s.rs, s.re, m.f[x]);
          mfmas(m.f);
        }
      });
  double* part = a.part + (size_t)sp * a.n * a.nb;
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = rbase + 16 * x + g + 4 * v, col = cbase + 16 * y + l16;
        if (r < a.n && col < a.nb) part[(size_t)r * a.nb + col] = s.w * acc[x][y][v];
      }
}

// Skinny block (nb <= 32 columns, e.g. the non-local projections of one
// atom set): wave w owns rows [32 w, 32 w + 32) of the 128-row block and
// both 16-column tiles, so no wave idles on absent columns.
__device__ __forceinline__ void gram_skinny_wave(const GramArgs& a, const double* __restrict__ B,
                                                 const GSlice& s, int sp, int bi) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, l16 = lane & 15, g = lane >> 4;
  const int rbase = bi * kGB + wave * 32;
  if (rbase >= a.n) return;
  int ra[2], rb[2];
#pragma unroll
  for (int x = 0; x < 2; x++) {  // rows past the matrices: outputs not stored
    const int r = rbase + 16 * x + l16, q = 16 * x + l16;
    ra[x] = r < a.n ? r : 0;
    rb[x] = q < a.nb ? q : 0;
  }
  v4d acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) acc[x][y] = vzero();
  auto mfmas = [&](const double (&f)[4][4]) __attribute__((always_inline)) {
#pragma unroll
    for (int e = 0; e < 4; e++)
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) acc[x][y] = mfma16(f[x][e], f[2 + y][e], acc[x][y]);
  };
  Frag<4> ring[kGramRing];  // fragments 0-1: rows of A, 2-3: rows of B
  chunk_ring(
      s.c0, s.c1, ring,
      [&](int c, Frag<4>& d) __attribute__((always_inline)) {
        const int kk = s.base + 16 * c + 4 * g;
#pragma unroll
        for (int x = 0; x < 2; x++) {
          gfrag(a.A, a.lda, ra[x], kk, d.f[x]);
          gfrag(B, a.lda, rb[x], kk, d.f[2 + x]);
        }
      },
      [&](int c, const Frag<4>& d) __attribute__((always_inline)) {
        const int k0 = s.base + 16 * c;
        if (k0 >= s.rs && k0 + 16 <= s.re) {
          mfmas(d.f);
        } else {  // a chunk at the region's edge (padding past K may hold anything)
          Frag<4> m = d;
#pragma unroll
          for (int x = 0; x < 4; x++) gmask(k0 + 4 * g, s.rs, s.re, m.f[x]);
          mfmas(m.f);
        }
      });
  double* part = a.part + (size_t)sp * a.n * a.nb;
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = rbase + 16 * x + g + 4 * v, col = 16 * y + l16;
        if (r < a.n && col < a.nb) part[(size_t)r * a.nb + col] = s.w * acc[x][y][v];
      }
}

// grid (output blocks, K slices); nbe = row blocks (blocks per edge when
// symmetric), nbe_b = column blocks, nlo = slices of the kx=0 region
__global__ __launch_bounds__(256) void k_gram(GramArgs a, int nbe, int nbe_b, int nlo) {
  int bi, bj, b = blockIdx.x;
  if (a.sym) {
    bi = 0;
    while (b >= nbe - bi) {
      b -= nbe - bi;
      bi++;
    }
    bj = bi + b;
  } else {
    bi = b / nbe_b;
    bj = b - bi * nbe_b;
  }
  const int sp = blockIdx.y;
  const GSlice s = gram_slice(a, sp, nlo);
  if (a.sym && bi == bj) {
    switch (threadIdx.x >> 6) {
      case 0: gram_diag_wave<0>(a, s, sp, bi); break;
      case 1: gram_diag_wave<1>(a, s, sp, bi); break;
      case 2: gram_diag_wave<2>(a, s, sp, bi); break;
      default: gram_diag_wave<3>(a, s, sp, bi); break;
    }
  } else if (!a.sym && a.nb <= 32) {
    gram_skinny_wave(a, a.B, s, sp, bi);
  } else {
    gram_full_wave(a, a.sym ? a.A : a.B, s, sp, bi, bj);
  }
}

int gram_slices(int n, int K, int k_lo, int sym, int target) {
  const int nbe = (n + kGB - 1) / kGB;
  const int blocks = sym ? nbe * (nbe + 1) / 2 : nbe * nbe;
  int ns = std::max(2, target / std::max(1, blocks));
  ns = std::min(ns, std::max(2, K / 64));  // at least four chunks per slice
  (void)k_lo;
  return ns;
}

hipError_t gram_f64(const GramArgs& a, hipStream_t st) {
  if (a.nslice < 1 || (a.k_lo > 0 && a.k_lo < a.K && a.nslice < 2) || (a.lda & 3) ||
      a.n < 1 || a.nb < 1 || (a.sym && a.nb != a.n))
    return hipErrorInvalidValue;
  const int nbe = (a.n + kGB - 1) / kGB, nbe_b = (a.nb + kGB - 1) / kGB;
  const int blocks = a.sym ? nbe * (nbe + 1) / 2 : nbe * nbe_b;
  hipLaunchKernelGGL(k_gram, dim3(blocks, a.nslice), dim3(256), 0, st, a, nbe, nbe_b,
                     gram_nlo(a));
  return hipGetLastError();
}

// C[e] = alpha * sum_sp part[sp][e'] with e' = e or its mirror; 64 outputs
// per workgroup, 16 groups of slices summed through LDS in a fixed order
template <typename P>
__global__ __launch_bounds__(1024) void k_gram_reduce(const P* __restrict__ part, int nslice,
                                                      int n, int nb, int sym, double alpha,
                                                      double* __restrict__ C) {
  __shared__ double red[16][64];
  const int x = threadIdx.x & 63, y = threadIdx.x >> 6;
  const size_t nn = (size_t)n * nb;
  const size_t e = (size_t)blockIdx.x * 64 + x;
  double s0 = 0.0, s1 = 0.0;
  if (e < nn) {
    int i = (int)(e / nb), j = (int)(e - (size_t)i * nb);
    if (sym && (i >> 4) > (j >> 4)) {  // sym: nb == n
      const int t = i;
      i = j;
      j = t;
    }
    const P* p = part + (size_t)i * nb + j;
    // thread y sums slices y, y+16, y+32, ...: 16 loads in flight,
    // branch-free (slice index clamped, value masked)
    for (int base = 0; base < nslice; base += 256) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int sp = base + y + 16 * u;
        const double w = (double)p[(size_t)min(sp, nslice - 1) * nn];
        v[u] = sp < nslice ? w : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 16; u += 2) {
        s0 += v[u];
        s1 += v[u + 1];
      }
    }
  }
  red[y][x] = s0 + s1;
  __syncthreads();
  if (y == 0 && e < nn) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < 16; w++) t += red[w][x];
    C[e] = alpha * t;
  }
}

// Symmetric product: four workgroups per upper 16 x 16 tile (the tiles the
// gram wrote), each a quarter of the tile's rows; written to C and, off the
// diagonal, to its mirror.  Thread (y, x): element x of the 4 x 16 quarter,
// slices y, y + 16, ...; loads branch-free (slice index clamped, value
// masked), 16 in flight; the 16 slice groups are summed through LDS in a
// fixed order.
template <typename P>
__global__ __launch_bounds__(1024) void k_gram_reduce_sym(const P* __restrict__ part,
                                                          int nslice, int n, double alpha,
                                                          double* __restrict__ C) {
  __shared__ double red[16][64];
  const int nt = (n + 15) >> 4;
  int b = blockIdx.x >> 2, ti = 0;
  while (b >= nt - ti) {
    b -= nt - ti;
    ti++;
  }
  const int tj = ti + b;
  const int x = threadIdx.x & 63, y = threadIdx.x >> 6;
  const int i = ti * 16 + (blockIdx.x & 3) * 4 + (x >> 4), j = tj * 16 + (x & 15);
  const size_t nn = (size_t)n * n;
  // a diagonal tile takes both halves from the upper triangle's partials, so
  // C is symmetric bit for bit also where the partials of (i, j) and (j, i)
  // were summed in different orders (split-BF16 term products)
  const int ir = ti == tj ? min(i, j) : i, jr = ti == tj ? max(i, j) : j;
  const P* p = part + (size_t)min(ir, n - 1) * n + min(jr, n - 1);
  double s0 = 0.0, s1 = 0.0;
  for (int base = 0; base < nslice; base += 256) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int sp = base + y + 16 * u;
      const double w = (double)p[(size_t)min(sp, nslice - 1) * nn];
      v[u] = sp < nslice ? w : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; u += 2) {
      s0 += v[u];
      s1 += v[u + 1];
    }
  }
  red[y][x] = s0 + s1;
  __syncthreads();
  if (y == 0 && i < n && j < n) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < 16; w++) t += red[w][x];
    const double v = alpha * t;
    C[(size_t)i * n + j] = v;
    if (ti < tj) C[(size_t)j * n + i] = v;  // a diagonal tile holds both halves
  }
}

template <typename P>
static hipError_t gram_reduce_t(const P* part, int nslice, int n, int nb, int sym, double alpha,
                                double* C, hipStream_t st) {
  if (sym && nb != n) return hipErrorInvalidValue;
  if (sym) {
    const int nt = (n + 15) >> 4;
    hipLaunchKernelGGL(k_gram_reduce_sym<P>, dim3(4 * (nt * (nt + 1) / 2)), dim3(1024), 0, st,
                       part, nslice, n, alpha, C);
    return hipGetLastError();
  }
  const size_t nn = (size_t)n * nb;
  hipLaunchKernelGGL(k_gram_reduce<P>, dim3((unsigned)((nn + 63) / 64)), dim3(1024), 0, st,
                     part, nslice, n, nb, sym, alpha, C);
  return hipGetLastError();
}

hipError_t gram_reduce(const double* part, int nslice, int n, int nb, int sym, double alpha,
                       double* C, hipStream_t st) {
  return gram_reduce_t(part, nslice, n, nb, sym, alpha, C, st);
}

// FP32 partials (split-BF16 gram), summed in FP64
hipError_t gram_reduce_f32(const float* part, int nslice, int n, int nb, int sym, double alpha,
                           double* C, hipStream_t st) {
  return gram_reduce_t(part, nslice, n, nb, sym, alpha, C, st);
}

// ================================================================== rotate
// Y[j][k] = beta Y[j][k] + s(k) sum_{i < ni} M[i][j] X[i][k] as D = M^T X: the
// output tile rows are j and its columns k, so the stores run along k.
// Workgroup: 64 columns k x 128 rows j; wave w owns j tiles 2w, 2w+1 of the
// block.  The i loop runs in chunks of 16 with the next chunk's operands
// loaded while the current chunk's 32 MFMAs issue.
constexpr int kRK = 64;

__global__ __launch_bounds__(256, 2) void k_rotate(RotArgs a) {
  __shared__ double red[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const int kb = blockIdx.x * kRK;
  const int jb = blockIdx.y * 128 + wave * 32;
  // operand columns clamped into range: columns past n or K are computed
  // from duplicated data and never stored
  int jc[2], kc[4];
#pragma unroll
  for (int x = 0; x < 2; x++) jc[x] = min(jb + 16 * x + l16, a.n - 1);
#pragma unroll
  for (int y = 0; y < 4; y++) kc[y] = min(kb + 16 * y + l16, a.K - 1);
  v4d acc[2][4];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) acc[x][y] = vzero();
  if (jb < a.n) {
    double am0[4][2], bx0[4][4], am1[4][2], bx1[4][4];
    // branch-free: rows past ni are clamped and their M values zeroed (a
    // load that only some lanes issue is waited for on its own)
    auto load = [&](int i0, double (&am)[4][2], double (&bx)[4][4]) {
#pragma unroll
      for (int s = 0; s < 4; s++) {
        const int i = i0 + 4 * s + g, ic = min(i, a.ni - 1);
        const bool vi = i < a.ni;
#pragma unroll
        for (int x = 0; x < 2; x++) {
          const double m = a.M[(size_t)ic * a.n + jc[x]];
          am[s][x] = vi ? m : 0.0;
        }
#pragma unroll
        for (int y = 0; y < 4; y++) bx[s][y] = a.X[(size_t)ic * a.ldx + kc[y]];
      }
    };
    auto mma = [&](const double (&am)[4][2], const double (&bx)[4][4]) {
#pragma unroll
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
          for (int y = 0; y < 4; y++) acc[x][y] = mfma16(am[s][x], bx[s][y], acc[x][y]);
    };
    int i0 = 0;
    load(0, am0, bx0);
    while (i0 < a.ni) {
      if (i0 + 16 < a.ni) load(i0 + 16, am1, bx1);
      mma(am0, bx0);
      i0 += 16;
      if (i0 >= a.ni) break;
      if (i0 + 16 < a.ni) load(i0 + 16, am0, bx0);
      mma(am1, bx1);
      i0 += 16;
    }
  }
  // fused steepest descent: every read of X (= sd_psi) in this workgroup
  // precedes the updates below; no other workgroup reads these columns
  if (a.sd_psi) __syncthreads();
  double fn = 0.0;
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 4; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int j = jb + 16 * x + g + 4 * v;
        const int k = kb + 16 * y + l16;  // unclamped: each output once
        if (j < a.n && k < a.K) {
          const double s = (k < a.k_lo) ? a.s_lo : a.s_hi;
          double* yp = a.Y + (size_t)j * a.ldx + k;
          double val = s * acc[x][y][v];
          if (a.beta != 0.0) val = a.beta * (*yp) + val;
          *yp = val;
          if (a.sd_psi) {
            a.sd_psi[(size_t)j * a.ldx + k] += a.dtm[k >> 1] * val;
            fn += val * val;
          }
        }
      }
  if (a.sd_psi) {
    for (int o = 32; o > 0; o >>= 1) fn += __shfl_down(fn, o, 64);
    if (lane == 0) red[wave] = fn;
    __syncthreads();
    if (tid == 0)
      a.fpart[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
  }
}

// n, ni <= 128: persistent workgroups of 8 waves (two per SIMD).  M is staged
// once per workgroup in LDS (all 128 x 128, zero beyond ni x n, row stride 129);
// workgroup b walks the 64-column k-blocks b, b + gridDim.x, ...  Wave w
// owns the output rows j in [32 (w & 3), +32) and the k tiles 2 (w >> 2),
// 2 (w >> 2) + 1 of a block.  The work items are (k-block, chunk of 16 i);
// a ring of R X buffers (R = 4, or 3 when the epilogue needs both Y and the
// SD operands) keeps the loads of item q+R in flight while the MFMAs of
// items q+1..q+R-1 issue, across k-block boundaries.  The
// epilogue's operands (Y when beta != 0, psi and dt/m for the SD update)
// are loaded a few items before the block ends, behind its MFMAs.
constexpr int rot_ring(bool use_y, bool use_sd) { return use_y && use_sd ? 3 : 4; }
constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }
constexpr int kRotLd = 129;                    // LDS row stride of M
constexpr int kRotThreads = 512;
constexpr size_t kRotSmem = (size_t)128 * kRotLd * sizeof(double);

struct XFrag {
  double v[4][2];  // [i step][k tile]
};

// NCH: i chunks of 16 per k-block (1, 2, 4 or 8; M is zero beyond ni).  The
// item loop is unrolled over G = lcm(NCH, R) items so that ring
// slots and block epilogues sit at fixed places of the loop body.  USE_Y
// (beta != 0) and USE_SD (sd_psi set) are compile-time so that the
// epilogue operands only take registers when used: the ring's loads must
// fit beside them to stay in flight.
template <int NCH, bool USE_Y, bool USE_SD>
__global__ __launch_bounds__(kRotThreads) void k_rotate_p(RotArgs a) {
  extern __shared__ __align__(16) double msm[];
  __shared__ double red[kRotThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  {  // stage M: every load of a thread is issued before its LDS stores
    constexpr int PER = 128 * 128 / kRotThreads;
    double mv[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int e = tid + kRotThreads * u, i = e >> 7, j = e & 127;
      const double m = a.M[(size_t)min(i, a.ni - 1) * a.n + min(j, a.n - 1)];
      mv[u] = (i < a.ni && j < a.n) ? m : 0.0;
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int e = tid + kRotThreads * u, i = e >> 7, j = e & 127;
      msm[i * kRotLd + j] = mv[u];
    }
  }
  __syncthreads();
  const int jb = (wave & 3) * 32, kt0 = 2 * (wave >> 2);
  const bool wact = jb < a.n;
  // Workgroup b owns the 16-column tiles [t0, t1), an even split of all
  // tiles (so at most one tile more than any other workgroup), taken in
  // blocks of four; tiles past t1 in a last partial block cost their waves
  // no MFMAs and are not stored.
  const int ntile = (a.K + 15) >> 4;
  const int t0 = (int)(((long long)blockIdx.x * ntile) / gridDim.x);
  const int t1 = (int)(((long long)(blockIdx.x + 1) * ntile) / gridDim.x);
  const int nmine = (t1 - t0 + 3) >> 2;
  constexpr int RING = rot_ring(USE_Y, USE_SD);
  constexpr int G = NCH / gcd_c(NCH, RING) * RING;
  const int nitems = nmine * NCH;
  auto kcol = [&](int q) { return (t0 + 4 * (q / NCH)) * 16; };  // first column of q's block
  // Branch-free loads (clamped to valid rows and columns) so that the
  // ring's loads stay in flight: rows i >= ni meet zero rows of the staged
  // M, columns k >= K produce outputs that are not stored.
  auto load = [&](int q, XFrag& d) __attribute__((always_inline)) {
    const int kb = kcol(q), ic = q % NCH;
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int i = min(ic * 16 + 4 * s + g, a.ni - 1);
#pragma unroll
      for (int y = 0; y < 2; y++) {
        const int k = min(kb + 16 * (kt0 + y) + l16, a.K - 1);
        d.v[s][y] = a.X[(size_t)i * a.ldx + k];
      }
    }
  };
  v4d acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) acc[x][y] = vzero();
  double yold[USE_Y ? 2 : 1][2][4], pold[USE_SD ? 2 : 1][2][4], dm[2];
  auto preload = [&](int kb) __attribute__((always_inline)) {
#pragma unroll
    for (int y = 0; y < 2; y++) {
      const int k = min(kb + 16 * (kt0 + y) + l16, a.K - 1);
      if (USE_SD) dm[y] = a.dtm[k >> 1];
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int j = min(jb + 16 * x + g + 4 * v, a.n - 1);
          if (USE_Y) yold[x][y][v] = a.Y[(size_t)j * a.ldx + k];
          if (USE_SD) pold[x][y][v] = a.sd_psi[(size_t)j * a.ldx + k];
        }
    }
  };
  double fn = 0.0;
  // output of one k-block: Y = beta Y + s(k) acc, optionally the SD update
  auto epilogue = [&](int kb) __attribute__((always_inline)) {
    if (USE_SD) __syncthreads();  // all waves are done reading these X columns
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 2; y++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int j = jb + 16 * x + g + 4 * v;
          const int k = kb + 16 * (kt0 + y) + l16;
          const double s = (k < a.k_lo) ? a.s_lo : a.s_hi;
          double val = s * acc[x][y][v];
          if (USE_Y) val = a.beta * yold[USE_Y ? x : 0][y][v] + val;
          if (j < a.n && k < a.K && (kb >> 4) + kt0 + y < t1) {
            a.Y[(size_t)j * a.ldx + k] = val;
            if (USE_SD)
              a.sd_psi[(size_t)j * a.ldx + k] = pold[USE_SD ? x : 0][y][v] + dm[y] * val;
            if (a.fpart) fn += val * val;
          }
          acc[x][y][v] = 0.0;
        }
  };
  XFrag ring[RING];
  if (nitems > 0) {
    // Every load below is issued on every path (items and blocks past the
    // workgroup's share load clamped addresses and are never stored): a
    // load that only some paths issue makes the compiler wait for all.
    static_for<0, RING>([&](auto dc) __attribute__((always_inline)) {
      constexpr int d = decltype(dc)::value;
      load(d, ring[d]);
    });
    // the epilogue's operands are loaded PRE items before the block ends
    // (short register lifetime, still behind MFMAs)
    constexpr int PRE = NCH >= 3 ? NCH - 3 : 0;
    for (int q0 = 0; q0 < nitems; q0 += G) {
      static_for<0, G>([&](auto dc) __attribute__((always_inline)) {
        constexpr int d = decltype(dc)::value;
        constexpr int sl = d % RING, ic = d % NCH;  // q0 is a multiple of G
        const int q = q0 + d;
        const int tb = t0 + 4 * (q / NCH) + kt0;  // this wave's first tile of the block
        const bool v0 = wact && tb < t1, v1 = wact && tb + 1 < t1;
        if (v0) {
#pragma unroll
          for (int s = 0; s < 4; s++) {
            const double* mrow = msm + (ic * 16 + 4 * s + g) * kRotLd + jb + l16;
            const double am0 = mrow[0], am1 = mrow[16];
            acc[0][0] = mfma16(am0, ring[sl].v[s][0], acc[0][0]);
            acc[1][0] = mfma16(am1, ring[sl].v[s][0], acc[1][0]);
            if (v1) {
              acc[0][1] = mfma16(am0, ring[sl].v[s][1], acc[0][1]);
              acc[1][1] = mfma16(am1, ring[sl].v[s][1], acc[1][1]);
            }
          }
        }
        if (ic == PRE) preload(kcol(q));
        load(q + RING, ring[sl]);
        if (ic == NCH - 1 && q < nitems) epilogue(kcol(q));
      });
    }
  }
  if (a.fpart) {
    for (int o = 32; o > 0; o >>= 1) fn += __shfl_down(fn, o, 64);
    if (lane == 0) red[wave] = fn;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < kRotThreads / 64; w++) t += red[w];
      a.fpart[blockIdx.x] = t;
    }
  }
}

static int rotate_pgrid(int K) {
  return std::max(1, std::min(cu_count(), (K + kRK - 1) / kRK));
}

int rotate_nwg(int n, int K) {
  return n <= 128 ? rotate_pgrid(K) : ((K + kRK - 1) / kRK) * ((n + 127) / 128);
}

hipError_t rotate_f64(const RotArgs& args, hipStream_t st) {
  RotArgs a = args;
  if (a.ni <= 0) a.ni = a.n;
  if ((a.sd_psi || a.fpart) && (a.n > 128 || a.ni > 128)) return hipErrorInvalidValue;
  if (a.sd_psi && (a.sd_psi != a.X || !a.dtm || !a.fpart || a.ni != a.n))
    return hipErrorInvalidValue;
  if (a.n <= 128 && a.ni <= 128) {
    const int nch = (a.ni + 15) >> 4;
    // [chunk-count variant][epilogue mode: Y read (beta != 0), SD update]
    const int v = 4 * (nch <= 1 ? 0 : nch <= 2 ? 1 : nch <= 4 ? 2 : 3) +
                  (a.beta != 0.0 ? 1 : 0) + (a.sd_psi ? 2 : 0);
#define OA_ROT(nc) k_rotate_p<nc, false, false>, k_rotate_p<nc, true, false>, \
                   k_rotate_p<nc, false, true>, k_rotate_p<nc, true, true>
    void (*const fns[16])(RotArgs) = {OA_ROT(1), OA_ROT(2), OA_ROT(4), OA_ROT(8)};
#undef OA_ROT
    if (const hipError_t e = grant_lds((const void*)fns[v], kRotSmem); e != hipSuccess) return e;
    hipLaunchKernelGGL(fns[v], dim3(rotate_pgrid(a.K)), dim3(kRotThreads), kRotSmem, st, a);
    return hipGetLastError();
  }
  dim3 grid((a.K + kRK - 1) / kRK, (a.n + 127) / 128);
  hipLaunchKernelGGL(k_rotate, grid, dim3(256), 0, st, a);
  return hipGetLastError();
}

// ================================================================== Newton-Schulz
// Workgroup = 32 x 32 tile of the n x n matrices, wave = 16 x 16 block.
#define NT 32
constexpr int kNsKP = 128;            // k panel width
constexpr int kNsLd = kNsKP + 1;      // LDS row stride of a row panel
constexpr int kNsCl = NT + 1;         // LDS row stride of a column panel
constexpr int kNsRowPanel = NT * kNsLd;
constexpr int kNsColPanel = kNsKP * kNsCl;
constexpr size_t kNsSmem = (size_t)2 * (kNsRowPanel + kNsColPanel) * sizeof(double);

// Stage A[r0.., k0..k0+KP) (row panel) and B[k0..k0+KP, c0..) (column panel)
// All of a thread's loads are issued before any LDS store, so a panel costs
// one memory latency rather than one per element.
constexpr int kNsPer = NT * kNsKP / 256;  // elements per thread and panel
struct NsPanel {
  double va[kNsPer], vb[kNsPer];
};
__device__ __forceinline__ void ns_load(const double* __restrict__ A,
                                        const double* __restrict__ B, int n, int r0, int c0,
                                        int k0, NsPanel& p) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int u = 0; u < kNsPer; u++) {
    const int e = tid + 256 * u;
    const int r = e / kNsKP, k = e - r * kNsKP;
    p.va[u] = (r0 + r < n && k0 + k < n) ? A[(size_t)(r0 + r) * n + k0 + k] : 0.0;
    const int kb = e / NT, cc = e - kb * NT;
    p.vb[u] = (k0 + kb < n && c0 + cc < n) ? B[(size_t)(k0 + kb) * n + c0 + cc] : 0.0;
  }
}
__device__ __forceinline__ void ns_store(const NsPanel& p, double* As, double* Bs) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int u = 0; u < kNsPer; u++) {
    const int e = tid + 256 * u;
    const int r = e / kNsKP, k = e - r * kNsKP;
    As[r * kNsLd + k] = p.va[u];
    const int kb = e / NT, cc = e - kb * NT;
    Bs[kb * kNsCl + cc] = p.vb[u];
  }
}

__device__ __forceinline__ v4d ns_mma(const double* As, const double* Bs, v4d acc) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, l16 = lane & 15, g = lane >> 4;
  const int wr = wave >> 1, wc = wave & 1;
  const double* ap = As + (16 * wr + l16) * kNsLd + g;
  const double* bp = Bs + g * kNsCl + 16 * wc + l16;
  // four independent accumulator chains keep the matrix pipe busy
  v4d p[4] = {acc, vzero(), vzero(), vzero()};
#pragma unroll
  for (int kk = 0; kk < kNsKP; kk += 16)
#pragma unroll
    for (int u = 0; u < 4; u++)
      p[u] = mfma16(ap[kk + 4 * u], bp[(kk + 4 * u) * kNsCl], p[u]);
  return (p[0] + p[1]) + (p[2] + p[3]);
}

// One (or two) tile products through LDS panels: acc1 = A1 B1 and, if A2,
// acc2 = A2 B2, for the workgroup's tile (r0, c0).
__device__ __forceinline__ void ns_tile2(const double* A1, const double* B1, const double* A2,
                                         const double* B2, int n, int r0, int c0, double* sm,
                                         v4d& acc1, v4d& acc2) {
  double* As1 = sm;
  double* Bs1 = As1 + kNsRowPanel;
  double* As2 = Bs1 + kNsColPanel;
  double* Bs2 = As2 + kNsRowPanel;
  acc1 = vzero();
  acc2 = vzero();
  for (int k0 = 0; k0 < n; k0 += kNsKP) {
    NsPanel p;
    ns_load(A1, B1, n, r0, c0, k0, p);
    __syncthreads();  // the previous panels are consumed
    ns_store(p, As1, Bs1);
    if (A2) {
      ns_load(A2, B2, n, r0, c0, k0, p);
      ns_store(p, As2, Bs2);
    }
    __syncthreads();
    acc1 = ns_mma(As1, Bs1, acc1);
    if (A2) acc2 = ns_mma(As2, Bs2, acc2);
  }
}

__device__ __forceinline__ void blk_coords(int nt, int& row0, int& col0) {
  const int wave = threadIdx.x >> 6;
  const int ti = blockIdx.x / nt, tj = blockIdx.x - (blockIdx.x / nt) * nt;
  row0 = ti * NT + (wave >> 1) * 16;
  col0 = tj * NT + (wave & 1) * 16;
}

// store alpha*acc + gamma*I; returns this lane's sum of (Told - new)^2
__device__ __forceinline__ double blk_store(double* __restrict__ C, const v4d& acc, int n,
                                            int row0, int col0, double alpha, double gamma,
                                            const double* __restrict__ Told) {
  const int lane = threadIdx.x & 63, l16 = lane & 15, g = lane >> 4;
  double e = 0.0;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int row = row0 + g + 4 * v, col = col0 + l16;
    if (row < n && col < n) {
      const double val = alpha * acc[v] + (row == col ? gamma : 0.0);
      C[(size_t)row * n + col] = val;
      if (Told) {
        const double d = Told[(size_t)row * n + col] - val;
        e += d * d;
      }
    }
  }
  return e;
}

__device__ __forceinline__ double wg_sum(double v) {
  __shared__ double red[4];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const double t = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return t;
}

// rms change of iteration `it` from the per-workgroup partials (same order
// in every workgroup, so every workgroup takes the same decision)
__device__ __forceinline__ double ns_err(const double* __restrict__ errpart, int it, int nwg,
                                         int n) {
  __shared__ double bcast;
  if (threadIdx.x == 0) {
    double s = 0.0;
    const double* ep = errpart + (size_t)it * nwg;
    for (int w = 0; w < nwg; w++) s += ep[w];
    bcast = sqrt(s / ((double)n * n));
  }
  __syncthreads();
  const double e = bcast;
  __syncthreads();
  return e;
}

// One launch for the whole iteration (cooperative, all workgroups resident):
// X0 = S/2, T0 = I; loop { C = 3I - T X; X' = X C / 2; T' = C T / 2 } with
// two grid barriers per iteration; stops when the rms change of T <= tol.
__global__ __launch_bounds__(256) void k_ns_coop(NsArgs a) {
  namespace cg = cooperative_groups;
  extern __shared__ __align__(16) double ns_smem[];
  cg::grid_group grid = cg::this_grid();
  const int n = a.n, nt = (n + NT - 1) / NT, nwg = nt * nt;
  const int r0 = (blockIdx.x / nt) * NT, c0 = (blockIdx.x % nt) * NT;
  int row0, col0;
  blk_coords(nt, row0, col0);
  const int lane = threadIdx.x & 63, l16 = lane & 15, g = lane >> 4;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int row = row0 + g + 4 * v, col = col0 + l16;
    if (row < n && col < n) {
      a.X0[(size_t)row * n + col] = 0.5 * a.S[(size_t)row * n + col];  // B = S/2 (ortho.C:236-239)
      a.T0[(size_t)row * n + col] = (row == col) ? 1.0 : 0.0;
    }
  }
  grid.sync();
  double* X[2] = {a.X0, a.X1};
  double* T[2] = {a.T0, a.T1};
  int conv = 0;
  double err = 0.0;
  for (int it = 1; it <= a.max_iter; it++) {
    const int o = (it - 1) & 1, w = it & 1;
    v4d p1, p2;
    ns_tile2(T[o], X[o], nullptr, nullptr, n, r0, c0, ns_smem, p1, p2);
    blk_store(a.C, p1, n, row0, col0, -1.0, 3.0, nullptr);
    grid.sync();
    ns_tile2(X[o], a.C, a.C, T[o], n, r0, c0, ns_smem, p1, p2);
    blk_store(X[w], p1, n, row0, col0, 0.5, 0.0, nullptr);
    const double e = blk_store(T[w], p2, n, row0, col0, 0.5, 0.0, T[o]);
    const double es = wg_sum(e);
    if (threadIdx.x == 0) a.errpart[(size_t)it * nwg + blockIdx.x] = es;
    grid.sync();
    err = ns_err(a.errpart, it, nwg, n);
    if (err <= a.tol) {
      conv = it;
      break;
    }
  }
  const int last = conv ? conv : a.max_iter;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.res->ns_iter = conv ? conv : -a.max_iter;
    a.res->ns_err = err;
  }
  const double* Tf = T[last & 1];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int row = row0 + g + 4 * v, col = col0 + l16;
    if (row < n && col < n) a.Tout[(size_t)row * n + col] = Tf[(size_t)row * n + col];
  }
}

// One workgroup for the whole iteration, a plain launch (n <= kNsOneMax):
// the recurrence of k_ns_coop with the workgroup visiting every 32 x 32 tile
// in turn and workgroup barriers in place of the grid barriers.  For the
// small matrices of small requests this replaces 2 max_iter + 2 launches per
// step (most of them early exits) by one, with no cooperative launch.
__global__ __launch_bounds__(256) void k_ns_one(NsArgs a) {
  extern __shared__ __align__(16) double ns_smem[];
  const int n = a.n, nt = (n + NT - 1) / NT, ntile = nt * nt;
  const int lane = threadIdx.x & 63, l16 = lane & 15, g = lane >> 4, wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * 16, wc = (wave & 1) * 16;
  for (int e = threadIdx.x; e < n * n; e += 256) {
    a.X0[e] = 0.5 * a.S[e];  // B = S/2 (ortho.C:236-239)
    a.T0[e] = (e / n == e % n) ? 1.0 : 0.0;
  }
  __syncthreads();
  double* X[2] = {a.X0, a.X1};
  double* T[2] = {a.T0, a.T1};
  int conv = 0;
  double err = 0.0;
  for (int it = 1; it <= a.max_iter; it++) {
    const int o = (it - 1) & 1, w = it & 1;
    v4d p1, p2;
    for (int t = 0; t < ntile; t++) {  // C = 3I - T X
      const int r0 = (t / nt) * NT, c0 = (t % nt) * NT;
      ns_tile2(T[o], X[o], nullptr, nullptr, n, r0, c0, ns_smem, p1, p2);
      blk_store(a.C, p1, n, r0 + wr, c0 + wc, -1.0, 3.0, nullptr);
    }
    __syncthreads();
    double e = 0.0;
    for (int t = 0; t < ntile; t++) {  // X' = X C / 2, T' = C T / 2
      const int r0 = (t / nt) * NT, c0 = (t % nt) * NT;
      ns_tile2(X[o], a.C, a.C, T[o], n, r0, c0, ns_smem, p1, p2);
      blk_store(X[w], p1, n, r0 + wr, c0 + wc, 0.5, 0.0, nullptr);
      e += blk_store(T[w], p2, n, r0 + wr, c0 + wc, 0.5, 0.0, T[o]);
    }
    err = sqrt(wg_sum(e) / ((double)n * n));  // wg_sum ends with a barrier
    if (err <= a.tol) {
      conv = it;
      break;
    }
  }
  const int last = conv ? conv : a.max_iter;
  if (threadIdx.x == 0) {
    a.res->ns_iter = conv ? conv : -a.max_iter;
    a.res->ns_err = err;
  }
  const double* Tf = T[last & 1];
  for (int e = threadIdx.x; e < n * n; e += 256) a.Tout[e] = Tf[e];
  (void)lane;
  (void)l16;
  (void)g;
}

// ---- fallback: two launches per iteration with a device-side early exit

// 16x16 block (row0, col0) of A B straight from global memory
__device__ __forceinline__ v4d blk_gemm(const double* __restrict__ A,
                                        const double* __restrict__ B, int n, int row0,
                                        int col0) {
  const int lane = threadIdx.x & 63, l16 = lane & 15, g = lane >> 4;
  const int r = row0 + l16, cc = col0 + l16;
  v4d acc = vzero();
  for (int k0 = 0; k0 < n; k0 += 4) {
    const int k = k0 + g;
    const double av = (r < n && k < n) ? A[(size_t)r * n + k] : 0.0;
    const double bv = (k < n && cc < n) ? B[(size_t)k * n + cc] : 0.0;
    acc = mfma16(av, bv, acc);
  }
  return acc;
}

__global__ void k_ns_init(const double* __restrict__ S, double* __restrict__ X0,
                          double* __restrict__ T0, int n, StepDev* res) {
  const size_t nn = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x) {
    X0[e] = S[e] * 0.5;
    const size_t i = e / n, j = e - i * n;
    T0[e] = (i == j) ? 1.0 : 0.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    res->ns_iter = 0;
    res->ns_err = 0.0;
  }
}

// FP64 iterations this launch sequence may run: max_iter, less the FP32
// iterations before it in the mixed-precision form
__device__ __forceinline__ int ns_limit(int max_iter, int mixed, const StepDev* res) {
  return mixed ? max_iter - *(volatile const int*)&res->ns32_iter : max_iter;
}

// True when iteration `it` must not run (converged before it, or past the
// limit).  Thread 0 decides and broadcasts, so the whole workgroup takes the
// same path.
__device__ bool ns_skip(const double* __restrict__ errpart, int it, int nwg, int n, double tol,
                        StepDev* res, bool record, int lim) {
  __shared__ int skip;
  if (threadIdx.x == 0) {
    int f = 0;
    if (it > lim || *(volatile int*)&res->ns_iter != 0) {
      f = 1;
    } else if (it > 1) {
      double s = 0.0;
      const double* ep = errpart + (size_t)(it - 1) * nwg;
      for (int w = 0; w < nwg; w++) s += ep[w];
      const double err = sqrt(s / ((double)n * n));
      if (err <= tol) {
        f = 1;
        if (record && blockIdx.x == 0) {
          res->ns_iter = it - 1;
          res->ns_err = err;
        }
      }
    }
    skip = f;
  }
  __syncthreads();
  return skip != 0;
}

__global__ __launch_bounds__(256) void k_ns_step1(const double* __restrict__ T,
                                                  const double* __restrict__ X,
                                                  double* __restrict__ C, int n, int it,
                                                  double tol, const double* errpart,
                                                  StepDev* res, int max_iter, int mixed) {
  const int nt = (n + NT - 1) / NT, nwg = nt * nt;
  if (ns_skip(errpart, it, nwg, n, tol, res, true, ns_limit(max_iter, mixed, res))) return;
  int row0, col0;
  blk_coords(nt, row0, col0);
  blk_store(C, blk_gemm(T, X, n, row0, col0), n, row0, col0, -1.0, 3.0, nullptr);
}

__global__ __launch_bounds__(256) void k_ns_step23(const double* __restrict__ Xo,
                                                   const double* __restrict__ To,
                                                   const double* __restrict__ C,
                                                   double* __restrict__ Xn,
                                                   double* __restrict__ Tn, int n, int it,
                                                   double tol, double* errpart, StepDev* res,
                                                   int max_iter, int mixed) {
  const int nt = (n + NT - 1) / NT, nwg = nt * nt;
  if (ns_skip(errpart, it, nwg, n, tol, res, false, ns_limit(max_iter, mixed, res))) return;
  int row0, col0;
  blk_coords(nt, row0, col0);
  blk_store(Xn, blk_gemm(Xo, C, n, row0, col0), n, row0, col0, 0.5, 0.0, nullptr);
  const double e = blk_store(Tn, blk_gemm(C, To, n, row0, col0), n, row0, col0, 0.5, 0.0, To);
  const double es = wg_sum(e);
  if (threadIdx.x == 0) errpart[(size_t)it * nwg + blockIdx.x] = es;
}

__global__ void k_ns_final(const double* __restrict__ errpart, int nwg, int n, int max_iter,
                           double tol, StepDev* res, int mixed) {
  if (threadIdx.x != 0) return;
  if (res->ns_iter != 0) return;
  const int lim = ns_limit(max_iter, mixed, res);
  double s = 0.0;
  const double* ep = errpart + (size_t)lim * nwg;
  for (int w = 0; w < nwg; w++) s += ep[w];
  const double err = sqrt(s / ((double)n * n));
  res->ns_err = err;
  res->ns_iter = (err <= tol) ? lim : -lim;
}

__global__ void k_ns_copy(const double* __restrict__ T0, const double* __restrict__ T1,
                          double* __restrict__ Tout, int n, const StepDev* res) {
  const int it = res->ns_iter < 0 ? -res->ns_iter : res->ns_iter;
  const double* src = (it & 1) ? T1 : T0;
  const size_t nn = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x)
    Tout[e] = src[e];
}

// X = (S/2) T in FP64, before each FP64 iteration of the mixed-precision
// form.  The coupled iteration keeps X = (S/2) T and sees S only through X,
// so it cannot remove an error of T that does not commute with S (an FP32
// T's): it converges to within that error (measured 1e-10 .. 1e-7).  With X
// rebuilt from the exact S, the iteration is T' = (3I - T (S/2) T) T / 2,
// which corrects such errors quadratically.
__global__ __launch_bounds__(256) void k_ns_x_from_t(const double* __restrict__ S,
                                                     const double* __restrict__ T,
                                                     double* __restrict__ X, int n, int it,
                                                     double tol, const double* errpart,
                                                     StepDev* res, int max_iter) {
  const int nt = (n + NT - 1) / NT, nwg = nt * nt;
  if (ns_skip(errpart, it, nwg, n, tol, res, false, ns_limit(max_iter, 1, res))) return;
  int row0, col0;
  blk_coords(nt, row0, col0);
  blk_store(X, blk_gemm(S, T, n, row0, col0), n, row0, col0, 0.5, 0.0, nullptr);
}

// init: X0 = S/2, T0 = I; else X0 and T0 are already set (the FP64 phase of
// the mixed-precision iteration)
static hipError_t ns_run_launches(const NsArgs& a, hipStream_t st, bool init = true,
                                  int mixed = 0) {
  const int n = a.n;
  const int nt = (n + NT - 1) / NT;
  const int nwg = nt * nt;
  const size_t nn = (size_t)n * n;
  int eb = (int)((nn + 255) / 256);
  if (eb > 512) eb = 512;
  if (init) hipLaunchKernelGGL(k_ns_init, dim3(eb), dim3(256), 0, st, a.S, a.X0, a.T0, n, a.res);
  double* X[2] = {a.X0, a.X1};
  double* T[2] = {a.T0, a.T1};
  for (int it = 1; it <= a.max_iter; it++) {
    const int o = (it - 1) & 1, w = it & 1;
    if (mixed)
      hipLaunchKernelGGL(k_ns_x_from_t, dim3(nwg), dim3(256), 0, st, a.S, T[o], X[o], n, it,
                         a.tol, a.errpart, a.res, a.max_iter);
    hipLaunchKernelGGL(k_ns_step1, dim3(nwg), dim3(256), 0, st, T[o], X[o], a.C, n, it, a.tol,
                       a.errpart, a.res, a.max_iter, mixed);
    hipLaunchKernelGGL(k_ns_step23, dim3(nwg), dim3(256), 0, st, X[o], T[o], a.C, X[w], T[w], n,
                       it, a.tol, a.errpart, a.res, a.max_iter, mixed);
  }
  hipLaunchKernelGGL(k_ns_final, dim3(1), dim3(64), 0, st, a.errpart, nwg, n, a.max_iter, a.tol,
                     a.res, mixed);
  hipLaunchKernelGGL(k_ns_copy, dim3(eb), dim3(256), 0, st, a.T0, a.T1, a.Tout, n, a.res);
  return hipGetLastError();
}

// ---- FP32 iterations of the mixed-precision Newton-Schulz
//
// The recurrence of get_iter_Tmat (cp_lowdin.C:201-241) on FP32 X, T, C with
// v_mfma_f32_16x16x4f32, two launches per iteration as above, until the rms
// change of T is <= kNs32Switch; the last FP32 T is then widened into the
// FP64 T0, and FP64 iterations with X = (S/2) T rebuilt from the exact S
// continue to the step's tolerance (k_ns_x_from_t): they remove the FP32
// error quadratically, and T meets the same tolerance as an FP64 run.
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f mfma16f(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// 16x16 block (row0, col0) of A B in FP32.  A lane holds A[r][k] and B[k][c]
// with r, c = lane % 16 and k = lane / 16, as in FP64; the FP32 accumulator
// element v of lane l is D[4 (l/16) + v][l % 16] (FP64: D[l/16 + 4 v]).
__device__ __forceinline__ v4f blk_gemm32(const float* __restrict__ A,
                                          const float* __restrict__ B, int n, int row0,
                                          int col0) {
  const int lane = threadIdx.x & 63, l16 = lane & 15, g = lane >> 4;
  const int r = row0 + l16, cc = col0 + l16;
  v4f acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < n; k0 += 4) {
    const int k = k0 + g;
    const float av = (r < n && k < n) ? A[(size_t)r * n + k] : 0.f;
    const float bv = (k < n && cc < n) ? B[(size_t)k * n + cc] : 0.f;
    acc = mfma16f(av, bv, acc);
  }
  return acc;
}

// C = alpha acc + gamma I on the block; returns sum (Told - C)^2 if Told
__device__ __forceinline__ double blk_store32(float* __restrict__ C, const v4f& acc, int n,
                                              int row0, int col0, float alpha, float gamma,
                                              const float* __restrict__ Told) {
  const int lane = threadIdx.x & 63, l16 = lane & 15, g = lane >> 4;
  double e = 0.0;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int row = row0 + 4 * g + v, col = col0 + l16;
    if (row < n && col < n) {
      const float val = alpha * acc[v] + (row == col ? gamma : 0.f);
      C[(size_t)row * n + col] = val;
      if (Told) {
        const double d = (double)Told[(size_t)row * n + col] - (double)val;
        e += d * d;
      }
    }
  }
  return e;
}

__global__ void k_ns32_init(const double* __restrict__ S, float* __restrict__ X0,
                            float* __restrict__ T0, int n, StepDev* res) {
  const size_t nn = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x) {
    X0[e] = (float)(S[e] * 0.5);
    const size_t i = e / n, j = e - i * n;
    T0[e] = (i == j) ? 1.f : 0.f;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) res->ns32_iter = 0;
}

// True when FP32 iteration `it` must not run: the switch test passed before
// it (recorded in ns32_iter by the first workgroup of step 1).
__device__ bool ns32_skip(const double* __restrict__ errpart, int it, int nwg, int n,
                          StepDev* res, bool record) {
  __shared__ int skip;
  if (threadIdx.x == 0) {
    int f = 0;
    if (*(volatile int*)&res->ns32_iter != 0) {
      f = 1;
    } else if (it > 1) {
      double s = 0.0;
      const double* ep = errpart + (size_t)(it - 1) * nwg;
      for (int w = 0; w < nwg; w++) s += ep[w];
      if (sqrt(s / ((double)n * n)) <= kNs32Switch) {
        f = 1;
        if (record && blockIdx.x == 0) res->ns32_iter = it - 1;
      }
    }
    skip = f;
  }
  __syncthreads();
  return skip != 0;
}

__global__ __launch_bounds__(256) void k_ns32_step1(const float* __restrict__ T,
                                                    const float* __restrict__ X,
                                                    float* __restrict__ C, int n, int it,
                                                    const double* errpart, StepDev* res) {
  const int nt = (n + NT - 1) / NT, nwg = nt * nt;
  if (ns32_skip(errpart, it, nwg, n, res, true)) return;
  int row0, col0;
  blk_coords(nt, row0, col0);
  blk_store32(C, blk_gemm32(T, X, n, row0, col0), n, row0, col0, -1.f, 3.f, nullptr);
}

__global__ __launch_bounds__(256) void k_ns32_step23(const float* __restrict__ Xo,
                                                     const float* __restrict__ To,
                                                     const float* __restrict__ C,
                                                     float* __restrict__ Xn,
                                                     float* __restrict__ Tn, int n, int it,
                                                     double* errpart, StepDev* res) {
  const int nt = (n + NT - 1) / NT, nwg = nt * nt;
  if (ns32_skip(errpart, it, nwg, n, res, false)) return;
  int row0, col0;
  blk_coords(nt, row0, col0);
  blk_store32(Xn, blk_gemm32(Xo, C, n, row0, col0), n, row0, col0, 0.5f, 0.f, nullptr);
  const double e =
      blk_store32(Tn, blk_gemm32(C, To, n, row0, col0), n, row0, col0, 0.5f, 0.f, To);
  const double es = wg_sum(e);
  if (threadIdx.x == 0) errpart[(size_t)it * nwg + blockIdx.x] = es;
}

// T of the last FP32 iterate (all max32 ran when none passed the switch
// test) widened into the FP64 T0
__global__ void k_ns32_widen(const float* __restrict__ T0f, const float* __restrict__ T1f,
                             double* __restrict__ T0, int n, int max_iter, StepDev* res) {
  int k = *(volatile int*)&res->ns32_iter;
  if (k == 0) k = max_iter;
  const float* T = (k & 1) ? T1f : T0f;
  const size_t nn = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x)
    T0[e] = (double)T[e];
}


__global__ void k_ns32_final(int max_iter, StepDev* res) {
  if (threadIdx.x == 0 && res->ns32_iter == 0) res->ns32_iter = max_iter;
  if (threadIdx.x == 0) {
    res->ns_iter = 0;
    res->ns_err = 0.0;
  }
}

hipError_t ns_run_mixed(const NsArgs& a, const Ns32Bufs& b, hipStream_t st) {
  const int n = a.n;
  const int nt = (n + NT - 1) / NT;
  const int nwg = nt * nt;
  const size_t nn = (size_t)n * n;
  int eb = (int)((nn + 255) / 256);
  if (eb > 512) eb = 512;
  hipLaunchKernelGGL(k_ns32_init, dim3(eb), dim3(256), 0, st, a.S, b.X0, b.T0, n, a.res);
  float* X[2] = {b.X0, b.X1};
  float* T[2] = {b.T0, b.T1};
  const int max32 = a.max_iter > 1 ? a.max_iter - 1 : 1;  // leaves at least one FP64 iteration
  for (int it = 1; it <= max32; it++) {
    const int o = (it - 1) & 1, w = it & 1;
    hipLaunchKernelGGL(k_ns32_step1, dim3(nwg), dim3(256), 0, st, T[o], X[o], b.C, n, it,
                       a.errpart, a.res);
    hipLaunchKernelGGL(k_ns32_step23, dim3(nwg), dim3(256), 0, st, X[o], T[o], b.C, X[w], T[w],
                       n, it, a.errpart, a.res);
  }
  // the widening reads ns32_iter before k_ns32_final sets "all ran"
  hipLaunchKernelGGL(k_ns32_widen, dim3(eb), dim3(256), 0, st, b.T0, b.T1, a.T0, n, max32,
                     a.res);
  hipLaunchKernelGGL(k_ns32_final, dim3(1), dim3(64), 0, st, max32, a.res);
  return ns_run_launches(a, st, false, 1);
}

// ---- block-row variant: one grid barrier per iteration
//
// Workgroup (b, h) owns rows R = [16b, 16b + 16) and the 16-column tiles
// ct = h, h + split, h + 2 split, ... of X and T (split = gridDim.y).  X, T
// and C are polynomials in S and commute, so the new X = X C / 2 of the
// recurrence equals C X / 2 (up to rounding), and the new tiles
//   X'[R,ct] = C[R,:] X[:,ct] / 2,  T'[R,ct] = C[R,:] T[:,ct] / 2
// need only C[R,:] = 3I - T[R,:] X, which every workgroup of the row block
// computes whole, and the previous X and T.  C[R,:] and T[R,:] sit in LDS;
// X and T go through global memory (double buffered) for the other
// workgroups.  Iteration 1 starts from T0 = I and X0 = S/2 (ortho.C:236-239):
// there C = 3I - S/2 and T1 = C/2 need no products, so no barrier precedes
// it.  After each barrier the next iteration's operands are loaded before
// the convergence test reads the error partials, so the two latencies
// overlap.  MFMA operands: A from LDS, B from global memory in batches of KB
// k-steps, branch-free (clamped addresses, masked values): a load that only
// some paths use is waited for on its own.
constexpr int kNrWaves = 8;
constexpr int kNrMaxWg = 256;  // error partials staged in LDS
// k-steps per B batch (two batches in registers at a time)
__host__ __device__ constexpr int nr_kb(int n) { return n <= 32 ? 8 : 16; }
// LDS row length: whole batches of k-steps (columns past n zero), + 2
// against bank conflicts
__host__ __device__ constexpr int nr_ld(int n) {
  return 4 * ((((n + 3) >> 2) + nr_kb(n) - 1) / nr_kb(n) * nr_kb(n)) + 2;
}
size_t ns_rows_smem(int n) { return (size_t)2 * 16 * nr_ld(n) * sizeof(double); }

// b[u] = scale * B[k][col] for the k-steps k0 .. k0 + KB - 1 of this lane
// group (k = 4 (k0 + u) + g): each load covers four rows of 16 contiguous
// columns.  k past n is clamped to the last row (finite values that meet
// zero A columns), col past n to the last column (results not stored): no
// per-k masks, which the compiler would hoist as scalar pairs and spill.
// Buffer loads: a scalar resource and one 32-bit offset register per load
// (n * n * 8 < 2^31), where flat addressing holds a register pair each.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
__device__ __forceinline__ rsrc_t nr_rsrc(const double* p, int n) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(p), (short)0, n * n * 8,
                                           0x00020000);  // raw buffer, gfx9 data format
}

template <int KB>
__device__ __forceinline__ void nr_load(double (&b)[KB], rsrc_t B, int col, int n, int k0,
                                        double scale) {
  const int g = (threadIdx.x & 63) >> 4;
  const int colc = min(col, n - 1);
#pragma unroll
  for (int u = 0; u < KB; u++) {
    const int k = min((k0 + u) * 4 + g, n - 1);
    const double v = __builtin_bit_cast(
        double, __builtin_amdgcn_raw_buffer_load_b64(B, (k * n + colc) * 8, 0, 0));
    b[u] = scale * v;
  }
}

// acc += A[16 LDS rows][k] b over one batch (A columns past n are zero)
template <int KB>
__device__ __forceinline__ v4d nr_mma(const double* __restrict__ Al, int ld, const double (&b)[KB],
                                      int k0, v4d acc) {
  const int lane = threadIdx.x & 63, l16 = lane & 15, g = lane >> 4;
  const double* __restrict__ a = Al + l16 * ld + 4 * k0 + g;
#pragma unroll
  for (int u = 0; u < KB; u++) acc = mfma16(a[4 * u], b[u], acc);
  return acc;
}

// A[16 LDS rows][0, n) B[0, n)[col tile]; b holds the first batch on entry
template <int KB>
__device__ __forceinline__ v4d nr_product(const double* __restrict__ Al, int ld, double (&b)[KB],
                                          rsrc_t B, int col, int n, double scale) {
  const int nk = (n + 3) >> 2;
  v4d acc = vzero();
#pragma nounroll
  for (int k0 = 0;; k0 += KB) {
    // the next batch is in flight while this one's MFMAs issue
    double nx[KB];
    const bool more = k0 + KB < nk;
    if (more) nr_load<KB>(nx, B, col, n, k0 + KB, scale);
    acc = nr_mma<KB>(Al, ld, b, k0, acc);
    if (!more) break;
#pragma unroll
    for (int u = 0; u < KB; u++) b[u] = nx[u];
  }
  return acc;
}

template <int KB>
__global__ __launch_bounds__(kNrWaves * 64) void k_ns_rows(NsArgs a) {
  namespace cg = cooperative_groups;
  static_assert(kNrWaves * 64 == 16 * 32, "A panel staging: 32 threads per row");
  extern __shared__ __align__(16) double nr_smem[];
  __shared__ double red[kNrWaves];
  __shared__ double epart[kNrMaxWg];
  __shared__ double bcast;
  cg::grid_group grid = cg::this_grid();
  const int nb = gridDim.x, split = gridDim.y;
  const int wg = blockIdx.y * nb + blockIdx.x, nwg = nb * split, h = blockIdx.y;
  const int r0 = 16 * blockIdx.x, tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63, l16 = lane & 15, g = lane >> 4;
  const int pr = tid >> 5, pk = tid & 31;  // A panel staging: row, first column
  constexpr int nthr = kNrWaves * 64;
  double bc[KB], bx[KB], ap[4];
  {
    const int n = a.n, ld = nr_ld(n);
    double* Tr = nr_smem;  // [16][ld]: rows R of T (previous iteration, then own tiles new)
    double* Cr = Tr + 16 * ld;  // [16][ld]: rows R of C
    nr_load<KB>(bx, nr_rsrc(a.S, n), (h + split * (wave >> 1)) * 16 + l16, n, 0, 0.5);  // X0 = S/2
    // iteration 1: T0 = I, C = 3I - S/2; LDS zero past n in both directions
    for (int e = tid; e < 16 * ld; e += nthr) {
      const int r = e / ld, k = e - r * ld, row = r0 + r;
      const bool in = row < n && k < n;
      const double sv = a.S[(size_t)min(row, n - 1) * n + min(k, n - 1)];
      Tr[e] = (in && row == k) ? 1.0 : 0.0;
      Cr[e] = in ? (row == k ? 3.0 : 0.0) - 0.5 * sv : 0.0;
    }
  }
  __syncthreads();
  int conv = 0;
  double err = 0.0;
  for (int it = 1; it <= a.max_iter; it++) {
    // n and what derives from it are recomputed in every iteration instead
    // of being hoisted into registers that the operand batches need
    int n = a.n;
    asm volatile("" : "+s"(n));
    const int ld = nr_ld(n), nct = (n + 15) >> 4, nown = (nct - h + split - 1) / split;
    double* Tr = nr_smem;
    double* Cr = Tr + 16 * ld;
    // X_k and T_k live in buffer k & 1
    const double* To = (it & 1) ? a.T0 : a.T1;
    double* Xw = (it & 1) ? a.X1 : a.X0;
    double* Tw = (it & 1) ? a.T1 : a.T0;
    const rsrc_t rXo = nr_rsrc((it & 1) ? a.X0 : a.X1, n), rTo = nr_rsrc(To, n);
    if (it > 1) {
      if (split > 1) {  // rows R of T_{it-1}: the other tiles come from other workgroups
        const bool rin = r0 + pr < n;
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (rin && pk + 32 * q < n) Tr[pr * ld + pk + 32 * q] = ap[q];
        for (int k = pk + 128; k < n; k += 32)
          if (rin) Tr[pr * ld + k] = To[(size_t)(r0 + pr) * n + k];
      }
      __syncthreads();
      // C[R,:] = 3I - T[R,:] X
#pragma nounroll
      for (int ct = wave, i = 0; ct < nct; ct += kNrWaves, i++) {
        const int col = ct * 16 + l16;
        if (i > 0) nr_load<KB>(bc, rXo, col, n, 0, 1.0);
        const v4d acc = nr_product<KB>(Tr, ld, bc, rXo, col, n, 1.0);
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int r = g + 4 * v, row = r0 + r;
          if (col < n) Cr[r * ld + col] = (row < n ? (row == col ? 3.0 : 0.0) - acc[v] : 0.0);
        }
      }
      __syncthreads();
    }
    // own tiles: X'[R,ct] = C[R,:] X[:,ct] / 2, T'[R,ct] = C[R,:] T[:,ct] / 2,
    // and the change of T
    double e = 0.0;
#pragma nounroll
    for (int q = wave, i = 0; q < 2 * nown; q += kNrWaves, i++) {
      const int m = q & 1, col = (h + split * (q >> 1)) * 16 + l16;
      v4d acc;
      if (it == 1 && m) {  // T0 = I: C T0 = C
#pragma unroll
        for (int v = 0; v < 4; v++) acc[v] = Cr[(g + 4 * v) * ld + min(col, ld - 1)];
      } else {
        const rsrc_t B = it == 1 ? nr_rsrc(a.S, n) : (m ? rTo : rXo);
        const double scale = it == 1 ? 0.5 : 1.0;
        if (i > 0) nr_load<KB>(bx, B, col, n, 0, scale);
        acc = nr_product<KB>(Cr, ld, bx, B, col, n, scale);
      }
      double* D = m ? Tw : Xw;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = g + 4 * v, row = r0 + r;
        if (row < n && col < n) {
          const double nv = 0.5 * acc[v];
          D[(unsigned)(row * n + col)] = nv;
          if (m) {
            const double d = Tr[r * ld + col] - nv;
            e += d * d;
            Tr[r * ld + col] = nv;  // this thread's element only: read before written
          }
        }
      }
    }
    {  // workgroup sum in a fixed order
      for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off, 64);
      if (lane == 0) red[wave] = e;
      __syncthreads();
      if (tid == 0) {
        double t = 0.0;
        for (int i = 0; i < kNrWaves; i++) t += red[i];
        a.errpart[(size_t)it * nwg + wg] = t;
      }
    }
    grid.sync();
    // the next iteration's first operands, then the error partials
    if (it < a.max_iter) {
      if (split > 1) {
        const double* p = Tw + (size_t)min(r0 + pr, n - 1) * n;
#pragma unroll
        for (int q = 0; q < 4; q++) ap[q] = p[min(pk + 32 * q, n - 1)];
      }
      const rsrc_t rXw = nr_rsrc(Xw, n);
      nr_load<KB>(bc, rXw, wave * 16 + l16, n, 0, 1.0);
      nr_load<KB>(bx, (wave & 1) ? nr_rsrc(Tw, n) : rXw, (h + split * (wave >> 1)) * 16 + l16, n,
                  0, 1.0);
    }
    for (int i = tid; i < nwg; i += nthr) epart[i] = a.errpart[(size_t)it * nwg + i];
    __syncthreads();
    if (tid == 0) {
      double sum = 0.0;
      for (int i = 0; i < nwg; i++) sum += epart[i];
      bcast = sqrt(sum / ((double)n * n));
    }
    __syncthreads();
    err = bcast;
    if (err <= a.tol) {
      conv = it;
      break;
    }
  }
  if (wg == 0 && tid == 0) {
    a.res->ns_iter = conv ? conv : -a.max_iter;
    a.res->ns_err = err;
  }
  // T of the last iteration: this workgroup's tiles are in its LDS rows
  const int n = a.n, ld = nr_ld(n), nct = (n + 15) >> 4, nown = (nct - h + split - 1) / split;
  for (int e = tid; e < 16 * nown * 16; e += nthr) {
    const int r = e / (16 * nown), t = e - r * 16 * nown;
    const int row = r0 + r, col = (h + split * (t >> 4)) * 16 + (t & 15);
    if (row < n && col < n) a.Tout[(size_t)row * n + col] = nr_smem[r * ld + col];
  }
}

static const void* nr_kernel(int n) {
  switch (nr_kb(n)) {
    case 8: return (const void*)k_ns_rows<8>;
    default: return (const void*)k_ns_rows<16>;
  }
}

static bool ns_smem_granted(int device) {
  int lds_max = 0;
  if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
          hipSuccess ||
      (size_t)lds_max < kNsSmem)
    return false;
  return grant_lds((const void*)k_ns_coop, kNsSmem) == hipSuccess;
}

// Can all n x n tiles be resident at once for the cooperative launch?
// The block-row variant: all ceil(n/16) x split workgroups resident, both
// LDS row blocks within the device's limit, at least one column tile per
// column group.
bool ns_rows_possible(int n, int split, int device) {
  int per_cu = 0, ncu = 0, coop = 0, lds_max = 0;
  const int nb = (n + 15) / 16;
  if (n < 1 || split < 1 || split > (n + 15) / 16 || nb * split > kNrMaxWg ||
      hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device) != hipSuccess ||
      !coop)
    return false;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
          hipSuccess)
    return false;
  const size_t smem = ns_rows_smem(n);
  const void* fn = nr_kernel(n);
  if (smem > (size_t)lds_max || grant_lds(fn, smem) != hipSuccess) return false;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kNrWaves * 64, smem) !=
      hipSuccess)
    return false;
  return (long)per_cu * ncu >= (long)nb * split;
}

bool ns_coop_possible(int n, int device) {
  const int nt = (n + NT - 1) / NT;
  int per_cu = 0, ncu = 0, coop = 0;
  if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device) != hipSuccess ||
      !coop)
    return false;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return false;
  if (!ns_smem_granted(device)) return false;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_ns_coop, 256,
                                                   kNsSmem) != hipSuccess)
    return false;
  return (long)per_cu * ncu >= (long)nt * nt;
}

hipError_t ns_run(const NsArgs& a, hipStream_t st) {
  if (!a.coop) return ns_run_launches(a, st);
  if (a.coop == kNsOneKind) {
    if (a.n > kNsOneMax) return ns_run_launches(a, st);
    const hipError_t e = grant_lds((const void*)k_ns_one, kNsSmem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_ns_one, dim3(1), dim3(256), kNsSmem, st, a);
    return hipGetLastError();
  }
  NsArgs args = a;
  void* params[] = {&args};
  hipError_t e;
  if (a.coop >= 2) {
    e = hipLaunchCooperativeKernel(nr_kernel(a.n), dim3((a.n + 15) / 16, ns_rows_split(a.coop)),
                                   dim3(kNrWaves * 64), params, ns_rows_smem(a.n), st);
  } else {
    const int nt = (a.n + NT - 1) / NT;
    e = hipLaunchCooperativeKernel((const void*)k_ns_coop, dim3(nt * nt), dim3(256), params,
                                   kNsSmem, st);
  }
  if (e == hipSuccess) return e;
  // The runtime refuses a cooperative launch when the device's grid-barrier
  // resource is held by another process (hipErrorIllegalState); a refused
  // launch enqueues nothing, so the same iteration runs as two launches per
  // iteration instead (its error partials fit the same buffer).
  (void)hipGetLastError();
  return ns_run_launches(a, st);
}

}  // namespace oa
