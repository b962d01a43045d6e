// oa_gemm_bf16.hip -- split-BF16 contractions of the mixed-precision step on
// CDNA4 matrix cores (v_mfma_f32_32x32x16_bf16).
//
// Every FP64 operand x is rounded to FP32, f = fl32(x) (2^-24 relative),
// and f is split exactly into three BF16 terms, f = h + m + l (h = bf16(f),
// m = bf16(f - h), l = f - h - m, each step exact in FP32 with round to
// nearest, |m| <= 2^-8 |f|, |l| <= 2^-16 |f|).  A product keeps the six term
// products with i + j <= 2 (h*h, h*m, m*h, h*l, m*m, l*h): each BF16 x BF16
// product is exact in FP32, the dropped ones are below 2^-24 relative, and
// the MFMA sums into FP32 accumulators.  So the result carries the FP32
// rounding of its operands and FP32 accumulation error over the products
// of one K slice (gram) or of the inner dimension (rotate), no more.
// K-slice partial sums are written in FP32 and summed in FP64.  The split
// costs three packed conversions, two widenings and two FP32 subtractions
// per pair of values (splitting the FP64 residuals instead took twice the
// vector instructions, which bound the gram).
//
//   gram_bf16   : C = sum_k w(k) A[i][k] B[j][k] (Lambda = Re Psi^H F, and S on
//                 request): one workgroup per 128 x 128 output block and K
//                 slice, wave (wr, wc) the 64 x 64 quadrant, operands loaded
//                 from HBM straight into registers and split there;
//   rotate_bf16 : Y = beta Y + s(k) M^T X (+ X) for n, ni <= 128 (F -= w Psi
//                 Lambda, and Psi T on request in the form Psi + Psi (T - I)):
//                 persistent workgroups over 64-column blocks of K; the
//                 split M^T (A operand) stays in registers for the whole
//                 launch, each block of X is split once into LDS in the
//                 B-operand layout and read back conflict-free.
//
// MFMA 32x32x16 lane layout (gfx950): A operand lane l holds row l%32 and
// k = 8 (l/32) + e, B operand lane l holds column l%32 and the same k; the
// accumulator element v of lane l is D[8 (v/4) + 4 (l/32) + v%4][l%32].
// Within one MFMA the assignment of k to (lane group, element) is free as
// long as A and B agree.
#include <hip/hip_runtime.h>

#include "oa_common.h"
#include "oa_gemm_common.h"
#include "oa_kernels.h"

namespace oa {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

namespace {

__device__ __forceinline__ f16v mfma32(bf8 a, bf8 b, f16v c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f16v fzero16() {
  f16v z;
#pragma unroll
  for (int v = 0; v < 16; v++) z[v] = 0.f;
  return z;
}

// Three BF16 terms of eight doubles (of their FP32 roundings).
struct Split {
  bf8 h, m, l;
};

__device__ __forceinline__ Split split3(const double (&x)[8]) {
  Split s;
  float r[8];
#pragma unroll
  for (int e = 0; e < 8; e++) r[e] = (float)x[e];
#pragma unroll
  for (int e = 0; e < 8; e++) s.h[e] = (__bf16)r[e];
#pragma unroll
  for (int e = 0; e < 8; e++) r[e] -= (float)s.h[e];
#pragma unroll
  for (int e = 0; e < 8; e++) s.m[e] = (__bf16)r[e];
#pragma unroll
  for (int e = 0; e < 8; e++) r[e] -= (float)s.m[e];
#pragma unroll
  for (int e = 0; e < 8; e++) s.l[e] = (__bf16)r[e];
  return s;
}

// acc += A B over the six kept term products, small terms first
__device__ __forceinline__ f16v mma6(const Split& a, const Split& b, f16v acc) {
  acc = mfma32(a.l, b.h, acc);
  acc = mfma32(a.m, b.m, acc);
  acc = mfma32(a.h, b.l, acc);
  acc = mfma32(a.m, b.h, acc);
  acc = mfma32(a.h, b.m, acc);
  return mfma32(a.h, b.h, acc);
}

// x[e] = X[row][kk + e], two 32-byte loads (kk a multiple of 8 inside the
// row's allocation; lda a multiple of 16)
__device__ __forceinline__ void load8(const double* __restrict__ X, int lda, int row, int kk,
                                      double (&x)[8]) {
  const double4* p = reinterpret_cast<const double4*>(X + (size_t)row * lda + kk);
  const double4 u = p[0], v = p[1];
  x[0] = u.x;
  x[1] = u.y;
  x[2] = u.z;
  x[3] = u.w;
  x[4] = v.x;
  x[5] = v.y;
  x[6] = v.z;
  x[7] = v.w;
}

__device__ __forceinline__ void mask8(int kk, int rs, int re, double (&x)[8]) {
#pragma unroll
  for (int e = 0; e < 8; e++) x[e] = (kk + e >= rs && kk + e < re) ? x[e] : 0.0;
}

struct GFrag {
  double a[2][8];  // two 32-row tiles of A
  double b[2][8];  // two 32-row tiles of B
};

}  // namespace

// grid (output blocks nbe x nbe_b, K slices); every block is computed whole
// (a symmetric product's lower half is mirrored by the reduction)
__global__ __launch_bounds__(256) void k_gram_bf16(GramArgs a, float* __restrict__ part,
                                                   int nbe_b, int nlo) {
  const int bi = blockIdx.x / nbe_b, bj = blockIdx.x - (blockIdx.x / nbe_b) * nbe_b;
  const int sp = blockIdx.y;
  const GSlice s = gram_slice(a, sp, nlo);
  const double* __restrict__ B = a.sym ? a.A : a.B;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l32 = lane & 31, g = lane >> 5;
  const int wr = wave >> 1, wc = wave & 1;
  const int rbase = bi * 128 + wr * 64, cbase = bj * 128 + wc * 64;
  if (rbase >= a.n || cbase >= a.nb) return;
  int ra[2], rb[2];
#pragma unroll
  for (int x = 0; x < 2; x++) {  // rows past the matrices: outputs not stored
    const int r = rbase + 32 * x + l32, q = cbase + 32 * x + l32;
    ra[x] = r < a.n ? r : 0;
    rb[x] = q < a.nb ? q : 0;
  }
  f16v acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) acc[x][y] = fzero16();
  auto mfmas = [&](const GFrag& f) __attribute__((always_inline)) {
    Split sa[2], sb[2];
#pragma unroll
    for (int x = 0; x < 2; x++) {
      sa[x] = split3(f.a[x]);
      sb[x] = split3(f.b[x]);
    }
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 2; y++) acc[x][y] = mma6(sa[x], sb[y], acc[x][y]);
  };
  GFrag ring[2];
  chunk_ring(
      s.c0, s.c1, ring,
      [&](int c, GFrag& d) __attribute__((always_inline)) {
        const int kk = s.base + 16 * c + 8 * g;
#pragma unroll
        for (int x = 0; x < 2; x++) {
          load8(a.A, a.lda, ra[x], kk, d.a[x]);
          load8(B, a.lda, rb[x], kk, d.b[x]);
        }
      },
      [&](int c, const GFrag& d) __attribute__((always_inline)) {
        const int k0 = s.base + 16 * c;
        if (k0 >= s.rs && k0 + 16 <= s.re) {
          mfmas(d);
        } else {  // a chunk at the region's edge (padding past K may hold anything)
          GFrag m = d;
#pragma unroll
          for (int x = 0; x < 2; x++) {
            mask8(k0 + 8 * g, s.rs, s.re, m.a[x]);
            mask8(k0 + 8 * g, s.rs, s.re, m.b[x]);
          }
          mfmas(m);
        }
      });
  float* out = part + (size_t)sp * a.n * a.nb;
  const float w = (float)s.w;  // 1 or 2: exact
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int r = rbase + 32 * x + 8 * (v >> 2) + 4 * g + (v & 3);
        const int col = cbase + 32 * y + l32;
        if (r < a.n && col < a.nb) out[(size_t)r * a.nb + col] = w * acc[x][y][v];
      }
}

hipError_t gram_bf16(const GramArgs& a, float* part, hipStream_t st) {
  if (a.nslice < 1 || (a.k_lo > 0 && a.k_lo < a.K && a.nslice < 2) || (a.lda & 15) ||
      a.n < 1 || a.nb < 1 || (a.sym && a.nb != a.n) || !part)
    return hipErrorInvalidValue;
  const int nbe = (a.n + 127) / 128, nbe_b = (a.nb + 127) / 128;
  hipLaunchKernelGGL(k_gram_bf16, dim3(nbe * nbe_b, a.nslice), dim3(256), 0, st, a, part, nbe_b,
                     gram_nlo(a));
  return hipGetLastError();
}

// ------------------------------------------------------------------ rotate
constexpr int kRbK = 64;  // columns of K per block
constexpr int kRbSteps = 8;  // inner-dimension steps of 16 (ni <= 128)
// LDS: [term][column tile][step][lane] of bf8 (16 bytes)
constexpr size_t kRbSmem = (size_t)3 * 2 * kRbSteps * 64 * sizeof(bf8);

__global__ __launch_bounds__(256, 1) void k_rotate_bf16(RotArgs a) {
  extern __shared__ __align__(16) unsigned char rb_smem[];
  bf8* Xs = reinterpret_cast<bf8*>(rb_smem);
  __shared__ double red[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l32 = lane & 31, g = lane >> 5;
  const int ni = a.ni, n = a.n;
  const int nsteps = (ni + 15) >> 4;
  // A operand: M^T split, rows j = 32 wave + l32, inner i = 16 s + 8 g + e
  // (M[i][j] at i*n + j; with m_minus_i, M - I)
  const int jrow = 32 * wave + l32;
  Split ms[kRbSteps];
#pragma unroll
  for (int s = 0; s < kRbSteps; s++) {
    double x[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int i = 16 * s + 8 * g + e;
      const bool ok = i < ni && jrow < n;
      const double v = a.M[(size_t)min(i, ni - 1) * n + min(jrow, n - 1)];
      x[e] = ok ? v - ((a.m_minus_i && i == jrow) ? 1.0 : 0.0) : 0.0;
    }
    ms[s] = split3(x);
  }
  const int nkb = (a.K + kRbK - 1) / kRbK;
  // staging: thread takes column kk = tid % 64 of a block and the 8-row
  // groups i8 = tid / 64 + 4 u of X; rows past ni are clamped (their M^T
  // entries are zero).  The next block's X is fetched into registers while
  // this block's products run.
  const int kk = tid & 63, w8 = tid >> 6;
  double xr[4][8];
  auto fetch = [&](int blk) __attribute__((always_inline)) {
    const int kb = min(blk, nkb - 1) * kRbK;
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int e = 0; e < 8; e++)
        xr[u][e] = a.X[(size_t)min(8 * (w8 + 4 * u) + e, ni - 1) * a.ldx + kb + kk];
  };
  fetch(blockIdx.x);
  double fsum = 0.0;
  for (int blk = blockIdx.x; blk < nkb; blk += gridDim.x) {
    const int kb = blk * kRbK;
    __syncthreads();  // the previous block's operands are consumed
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int i8 = w8 + 4 * u;
      const Split sx = split3(xr[u]);
      const int idx = ((kk >> 5) * kRbSteps + (i8 >> 1)) * 64 + (kk & 31) + 32 * (i8 & 1);
      Xs[idx] = sx.h;
      Xs[2 * kRbSteps * 64 + idx] = sx.m;
      Xs[4 * kRbSteps * 64 + idx] = sx.l;
    }
    // what the epilogue adds (beta Y + Xadd), loaded unconditionally at
    // clamped addresses; only the stores are masked
    double old[2][16];
#pragma unroll
    for (int kt = 0; kt < 2; kt++) {
      const int kc = min(kb + 32 * kt + l32, a.K - 1);
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int jc = min(32 * wave + 8 * (v >> 2) + 4 * g + (v & 3), n - 1);
        const size_t o = (size_t)jc * a.ldx + kc;
        old[kt][v] = (a.beta != 0.0 ? a.beta * a.Y[o] : 0.0) + (a.Xadd ? a.Xadd[o] : 0.0);
      }
    }
    fetch(blk + gridDim.x);
    __syncthreads();
    f16v acc[2] = {fzero16(), fzero16()};
#pragma unroll
    for (int s = 0; s < kRbSteps; s++) {
      if (s < nsteps) {
#pragma unroll
        for (int kt = 0; kt < 2; kt++) {
          const int idx = (kt * kRbSteps + s) * 64 + lane;
          Split xb;
          xb.h = Xs[idx];
          xb.m = Xs[2 * kRbSteps * 64 + idx];
          xb.l = Xs[4 * kRbSteps * 64 + idx];
          acc[kt] = mma6(ms[s], xb, acc[kt]);
        }
      }
    }
    // epilogue: rows j, columns k along the lanes (256-byte stores)
#pragma unroll
    for (int kt = 0; kt < 2; kt++) {
      const int k = kb + 32 * kt + l32;
      const double sk = k < a.k_lo ? a.s_lo : a.s_hi;
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int j = 32 * wave + 8 * (v >> 2) + 4 * g + (v & 3);
        const double y = sk * (double)acc[kt][v] + old[kt][v];
        if (j < n && k < a.K) {
          a.Y[(size_t)j * a.ldx + k] = y;
          fsum += y * y;
        }
      }
    }
  }
  if (a.fpart) {
    for (int o = 32; o > 0; o >>= 1) fsum += __shfl_down(fsum, o, 64);
    if (lane == 0) red[wave] = fsum;
    __syncthreads();
    if (tid == 0) a.fpart[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
  }
}

int rotate_bf16_nwg(int K) {
  return std::max(1, std::min(cu_count(), (K + kRbK - 1) / kRbK));
}

hipError_t rotate_bf16(const RotArgs& args, hipStream_t st) {
  RotArgs a = args;
  if (a.ni <= 0) a.ni = a.n;
  if (a.n > 128 || a.ni > 128 || a.sd_psi || (a.ldx % kRbK) || a.K > a.ldx || a.Xadd == a.Y)
    return hipErrorInvalidValue;
  if (const hipError_t e = grant_lds((const void*)k_rotate_bf16, kRbSmem); e != hipSuccess)
    return e;
  hipLaunchKernelGGL(k_rotate_bf16, dim3(rotate_bf16_nwg(a.K)), dim3(256), kRbSmem, st, a);
  return hipGetLastError();
}

}  // namespace oa
