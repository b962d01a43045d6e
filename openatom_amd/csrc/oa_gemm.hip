// oa_gemm.hip -- dense FP64 contractions of the CP electronic step on
// CDNA4 matrix cores (v_mfma_f64_16x16x4f64), plus the small pointwise
// G-space kernels.
//
//   gram    : S = 2 Re Psi^H Psi over stored points with kx=0 halved
//             (PairCalculator::multiplyForward, ckPairCalculator.C:682-802)
//             and Lambda = Re Psi^H F (asymmetric PairCalculator, alpha = 1);
//             split-K over the long point axis, symmetric case computes the
//             upper tiles only.
//   rotate  : Psi' = Psi T and F -= w Psi Lambda
//             (bwMultiplyHelper, ckPairCalculator.C:1189-1425;
//              unpackLambda, CP_State_GSpacePlane.C:2045-2225).
//   ns_run  : Newton-Schulz iteration for (S/2)^{-1/2}
//             (Ortho::do_iteration/step_2/step_3/tolerance_check,
//              ortho.C:862-985; serial twin cp_lowdin.C:180-265).
#include "oa_common.h"
#include "oa_kernels.h"

namespace oa {

typedef double v4d __attribute__((ext_vector_type(4)));

// D(16x16) += A(16x4) B(4x16).  Lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; it receives D[l/16 + 4*v][l%16] in element v (layout
// measured on gfx950 with a single-wave probe).
__device__ __forceinline__ v4d mfma16(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------------ gram
#define GT 64
#define GKC 16

__global__ __launch_bounds__(256) void k_gram(GramArgs a, int ntile) {
  __shared__ double As[GT][GKC + 1];
  __shared__ double Bs[GT][GKC + 1];
  int ti, tj;
  if (a.sym) {
    // blockIdx.x enumerates upper-triangle tile pairs row by row
    int b = blockIdx.x;
    ti = 0;
    while (b >= ntile - ti) { b -= ntile - ti; ti++; }
    tj = ti + b;
  } else {
    ti = blockIdx.x / ntile;
    tj = blockIdx.x - ti * ntile;
  }
  const int sp = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = wave >> 1, wj = wave & 1;
  const int l16 = lane & 15, lq = lane >> 4;
  v4d acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) acc[x][y] = (v4d){0.0, 0.0, 0.0, 0.0};
  const int kbeg = sp * a.kslice;
  const int kend = min(a.K, kbeg + a.kslice);
  const double* B = a.sym ? a.A : a.B;
  for (int k0 = kbeg; k0 < kend; k0 += GKC) {
    for (int e = tid; e < GT * GKC; e += 256) {
      const int r = e / GKC, kk = e - r * GKC, k = k0 + kk;
      const int i = ti * GT + r, j = tj * GT + r;
      const double w = (k < a.k_lo) ? a.w_lo : a.w_hi;
      As[r][kk] = (i < a.n && k < kend) ? a.A[(size_t)i * a.lda + k] * w : 0.0;
      Bs[r][kk] = (j < a.n && k < kend) ? B[(size_t)j * a.lda + k] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kq = 0; kq < GKC; kq += 4) {
      double av[2], bv[2];
#pragma unroll
      for (int x = 0; x < 2; x++) av[x] = As[wi * 32 + x * 16 + l16][kq + lq];
#pragma unroll
      for (int y = 0; y < 2; y++) bv[y] = Bs[wj * 32 + y * 16 + l16][kq + lq];
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) acc[x][y] = mfma16(av[x], bv[y], acc[x][y]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = ti * GT + wi * 32 + x * 16 + lq + 4 * v;
        const int col = tj * GT + wj * 32 + y * 16 + l16;
        if (row < a.n && col < a.n) a.part[((size_t)sp * a.n + row) * a.n + col] = acc[x][y][v];
      }
}

int gram_nsplit(int K, int kslice) { return (K + kslice - 1) / kslice; }

hipError_t gram_f64(const GramArgs& a, hipStream_t st) {
  const int ntile = (a.n + GT - 1) / GT;
  const int ntiles = a.sym ? ntile * (ntile + 1) / 2 : ntile * ntile;
  dim3 grid(ntiles, gram_nsplit(a.K, a.kslice));
  hipLaunchKernelGGL(k_gram, grid, dim3(256), 0, st, a, ntile);
  return hipGetLastError();
}

__global__ void k_gram_reduce(const double* __restrict__ part, int nsplit, int n, int sym,
                              double alpha, double* __restrict__ C) {
  const size_t nn = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(e / n), j = (int)(e - (size_t)i * n);
    int si = i, sj = j;
    if (sym && (i / GT) > (j / GT)) { si = j; sj = i; }
    double acc = 0.0;
    for (int sp = 0; sp < nsplit; sp++) acc += part[((size_t)sp * n + si) * n + sj];
    C[e] = alpha * acc;
  }
}

hipError_t gram_reduce(const double* part, int nsplit, int n, int sym, double alpha, double* C,
                       hipStream_t st) {
  const size_t nn = (size_t)n * n;
  int blocks = (int)((nn + 255) / 256);
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_gram_reduce, dim3(blocks), dim3(256), 0, st, part, nsplit, n, sym, alpha,
                     C);
  return hipGetLastError();
}

// ------------------------------------------------------------------ rotate
#define RT 64
#define RKC 16

__global__ __launch_bounds__(256) void k_rotate(RotArgs a) {
  __shared__ double Ms[RT][RKC + 1];  // Ms[jl][il] = M[i0+il][j0+jl]
  __shared__ double Xs[RKC][RT + 1];  // Xs[il][kl] = X[i0+il][k0+kl]
  const int k0 = blockIdx.x * RT, j0 = blockIdx.y * RT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wj = wave >> 1, wk = wave & 1;
  const int l16 = lane & 15, lq = lane >> 4;
  v4d acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) acc[x][y] = (v4d){0.0, 0.0, 0.0, 0.0};
  for (int i0 = 0; i0 < a.n; i0 += RKC) {
    for (int e = tid; e < RT * RKC; e += 256) {
      const int il = e / RT, jl = e - il * RT;
      const int i = i0 + il;
      Ms[jl][il] = (i < a.n && j0 + jl < a.n) ? a.M[(size_t)i * a.n + j0 + jl] : 0.0;
      Xs[il][jl] = (i < a.n && k0 + jl < a.K) ? a.X[(size_t)i * a.ldx + k0 + jl] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int iq = 0; iq < RKC; iq += 4) {
      double av[2], bv[2];
#pragma unroll
      for (int x = 0; x < 2; x++) av[x] = Ms[wj * 32 + x * 16 + l16][iq + lq];
#pragma unroll
      for (int y = 0; y < 2; y++) bv[y] = Xs[iq + lq][wk * 32 + y * 16 + l16];
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) acc[x][y] = mfma16(av[x], bv[y], acc[x][y]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int j = j0 + wj * 32 + x * 16 + lq + 4 * v;
        const int k = k0 + wk * 32 + y * 16 + l16;
        if (j < a.n && k < a.K) {
          const double s = (k < a.k_lo) ? a.s_lo : a.s_hi;
          double* yp = a.Y + (size_t)j * a.ldx + k;
          const double val = s * acc[x][y][v];
          *yp = (a.beta != 0.0) ? a.beta * (*yp) + val : val;
        }
      }
}

hipError_t rotate_f64(const RotArgs& a, hipStream_t st) {
  dim3 grid((a.K + RT - 1) / RT, (a.n + RT - 1) / RT);
  hipLaunchKernelGGL(k_rotate, grid, dim3(256), 0, st, a);
  return hipGetLastError();
}

// ------------------------------------------------------------------ Newton-Schulz
#define NT 32

// C(tile) = alpha * A B + beta * D + gamma * I over a 32x32 tile; four waves
// each own a 16x16 block.  Returns this workgroup's sum of (Told - C)^2 when
// Told is given.
__device__ double small_gemm_tile(const double* __restrict__ A, const double* __restrict__ B,
                                  double* __restrict__ C, int n, double alpha, double gamma,
                                  const double* __restrict__ Told, int ti, int tj) {
  __shared__ double As[NT][17];
  __shared__ double Bs[16][NT + 1];
  __shared__ double red[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = wave >> 1, wj = wave & 1;
  const int l16 = lane & 15, lq = lane >> 4;
  v4d acc = (v4d){0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < n; k0 += 16) {
    for (int e = tid; e < NT * 16; e += 256) {
      const int r = e / 16, kk = e - r * 16;
      const int i = ti * NT + r, k = k0 + kk;
      As[r][kk] = (i < n && k < n) ? A[(size_t)i * n + k] : 0.0;
      const int kb = e / NT, c = e - kb * NT;
      const int kr = k0 + kb, j = tj * NT + c;
      Bs[kb][c] = (kr < n && j < n) ? B[(size_t)kr * n + j] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kq = 0; kq < 16; kq += 4)
      acc = mfma16(As[wi * 16 + l16][kq + lq], Bs[kq + lq][wj * 16 + l16], acc);
    __syncthreads();
  }
  double err = 0.0;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int row = ti * NT + wi * 16 + lq + 4 * v;
    const int col = tj * NT + wj * 16 + l16;
    if (row < n && col < n) {
      double val = alpha * acc[v] + (row == col ? gamma : 0.0);
      C[(size_t)row * n + col] = val;
      if (Told) {
        const double d = Told[(size_t)row * n + col] - val;
        err += d * d;
      }
    }
  }
  if (!Told) return 0.0;
  for (int o = 32; o > 0; o >>= 1) err += __shfl_down(err, o, 64);
  if (lane == 0) red[wave] = err;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

__global__ void k_ns_init(const double* __restrict__ S, double* __restrict__ X0,
                          double* __restrict__ T0, int n, StepDev* res) {
  const size_t nn = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x) {
    X0[e] = S[e] * 0.5;  // B = S/2 (ortho.C:236-239), occupation 2 folded in
    const size_t i = e / n, j = e - i * n;
    T0[e] = (i == j) ? 1.0 : 0.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    res->ns_iter = 0;
    res->ns_err = 0.0;
  }
}

// Returns true when iteration `it` must not run (already converged).
__device__ bool ns_skip(const double* __restrict__ errpart, int it, int nwg, int n, double tol,
                        StepDev* res, bool record) {
  if (*(volatile int*)&res->ns_iter != 0) return true;
  if (it <= 1) return false;
  double s = 0.0;
  const double* ep = errpart + (size_t)(it - 1) * nwg;
  for (int w = 0; w < nwg; w++) s += ep[w];
  const double err = sqrt(s / ((double)n * n));
  if (err <= tol) {
    if (record && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
      res->ns_iter = it - 1;
      res->ns_err = err;
    }
    return true;
  }
  return false;
}

// step 1: C = 3I - T X (ortho.C:859-879)
__global__ __launch_bounds__(256) void k_ns_step1(const double* __restrict__ T,
                                                  const double* __restrict__ X,
                                                  double* __restrict__ C, int n, int it,
                                                  double tol, const double* errpart,
                                                  StepDev* res) {
  const int nwg = gridDim.x * gridDim.y;
  if (ns_skip(errpart, it, nwg, n, tol, res, true)) return;
  small_gemm_tile(T, X, C, n, -1.0, 3.0, nullptr, blockIdx.x, blockIdx.y);
}

// steps 2 and 3: X' = 1/2 X C ; T' = 1/2 C T, with the rms change of T
// (ortho.C:894-985); blockIdx.z selects the product.
__global__ __launch_bounds__(256) void k_ns_step23(const double* __restrict__ Xo,
                                                   const double* __restrict__ To,
                                                   const double* __restrict__ C,
                                                   double* __restrict__ Xn,
                                                   double* __restrict__ Tn, int n, int it,
                                                   double tol, double* errpart, StepDev* res) {
  const int nwg = gridDim.x * gridDim.y;
  if (ns_skip(errpart, it, nwg, n, tol, res, false)) return;
  if (blockIdx.z == 0) {
    small_gemm_tile(Xo, C, Xn, n, 0.5, 0.0, nullptr, blockIdx.x, blockIdx.y);
  } else {
    const double e = small_gemm_tile(C, To, Tn, n, 0.5, 0.0, To, blockIdx.x, blockIdx.y);
    if (threadIdx.x == 0) errpart[(size_t)it * nwg + blockIdx.y * gridDim.x + blockIdx.x] = e;
  }
}

__global__ void k_ns_final(const double* __restrict__ errpart, int nwg, int n, int max_iter,
                           double tol, StepDev* res) {
  if (threadIdx.x != 0) return;
  if (res->ns_iter != 0) return;
  double s = 0.0;
  const double* ep = errpart + (size_t)max_iter * nwg;
  for (int w = 0; w < nwg; w++) s += ep[w];
  const double err = sqrt(s / ((double)n * n));
  res->ns_err = err;
  res->ns_iter = (err <= tol) ? max_iter : -max_iter;
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

hipError_t ns_run(const NsArgs& a, hipStream_t st) {
  const int n = a.n;
  const int nt = (n + NT - 1) / NT;
  const int nwg = nt * nt;
  const size_t nn = (size_t)n * n;
  int eb = (int)((nn + 255) / 256);
  if (eb > 512) eb = 512;
  hipLaunchKernelGGL(k_ns_init, dim3(eb), dim3(256), 0, st, a.S, a.X0, a.T0, n, a.res);
  double* X[2] = {a.X0, a.X1};
  double* T[2] = {a.T0, a.T1};
  for (int it = 1; it <= a.max_iter; it++) {
    const int o = (it - 1) & 1, w = it & 1;
    hipLaunchKernelGGL(k_ns_step1, dim3(nt, nt), dim3(256), 0, st, T[o], X[o], a.C, n, it,
                       a.tol, a.errpart, a.res);
    hipLaunchKernelGGL(k_ns_step23, dim3(nt, nt, 2), dim3(256), 0, st, X[o], T[o], a.C, X[w],
                       T[w], n, it, a.tol, a.errpart, a.res);
  }
  hipLaunchKernelGGL(k_ns_final, dim3(1), dim3(64), 0, st, a.errpart, nwg, n, a.max_iter, a.tol,
                     a.res);
  hipLaunchKernelGGL(k_ns_copy, dim3(eb), dim3(256), 0, st, a.T0, a.T1, a.Tout, n, a.res);
  return hipGetLastError();
}

// ------------------------------------------------------------------ pointwise

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[16];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
  return t;
}

__global__ void k_sd(double2* __restrict__ psi, const double2* __restrict__ f,
                     const double* __restrict__ dtm, int nst, int npts, int ld,
                     double* __restrict__ fpart) {
  const size_t tot = (size_t)nst * npts;
  double fn = 0.0;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(e / npts), p = (int)(e - (size_t)s * npts);
    const size_t o = (size_t)s * ld + p;
    const double2 fv = f[o];
    const double m = dtm[p];
    double2 c = psi[o];
    c.x += fv.x * m;
    c.y += fv.y * m;
    psi[o] = c;
    fn += fv.x * fv.x + fv.y * fv.y;
  }
  const double t = block_sum(fn);
  if (threadIdx.x == 0) fpart[blockIdx.x] = t;
}

hipError_t sd_update(double2* psi, const double2* f, const double* dtm, int nst, int npts, int ld,
                     double* fpart, int nfpart, hipStream_t st) {
  hipLaunchKernelGGL(k_sd, dim3(nfpart), dim3(256), 0, st, psi, f, dtm, nst, npts, ld, fpart);
  return hipGetLastError();
}

__global__ void k_red_sync(double2* __restrict__ psi, const int* __restrict__ red,
                           const int* __restrict__ partner, int nred, int nst, int ld) {
  const size_t tot = (size_t)nst * nred;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(e / nred), r = (int)(e - (size_t)s * nred);
    const double2 v = psi[(size_t)s * ld + partner[r]];
    psi[(size_t)s * ld + red[r]] = make_double2(v.x, -v.y);
  }
}

hipError_t red_sync(double2* psi, const int* red, const int* partner, int nred, int nst, int ld,
                    hipStream_t st) {
  if (nred == 0) return hipSuccess;
  const size_t tot = (size_t)nst * nred;
  int blocks = (int)((tot + 255) / 256);
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_red_sync, dim3(blocks), dim3(256), 0, st, psi, red, partner, nred, nst,
                     ld);
  return hipGetLastError();
}

__global__ void k_reduce_to(const double* __restrict__ part, int n, double scale,
                            double* __restrict__ out) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v += part[i];
  const double t = block_sum(v);
  if (threadIdx.x == 0) *out = t * scale;
}

hipError_t reduce_to(const double* part, int n, double scale, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_to, dim3(1), dim3(1024), 0, st, part, n, scale, out);
  return hipGetLastError();
}

__global__ void k_dot(const double* __restrict__ a, const double* __restrict__ b, size_t n,
                      double* __restrict__ part) {
  double v = 0.0;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x)
    v += a[e] * b[e];
  const double t = block_sum(v);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

hipError_t dot_parts(const double* a, const double* b, size_t n, double* part, int nparts,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_dot, dim3(nparts), dim3(256), 0, st, a, b, n, part);
  return hipGetLastError();
}

__global__ void k_resid(const double* __restrict__ S, int n, StepDev* res) {
  __shared__ double rmax[16];
  double mx = 0.0, ss = 0.0;
  const size_t nn = (size_t)n * n;
  for (size_t e = threadIdx.x; e < nn; e += blockDim.x) {
    const size_t i = e / n, j = e - i * n;
    const double d = 0.5 * S[e] - (i == j ? 1.0 : 0.0);
    mx = fmax(mx, fabs(d));
    ss += d * d;
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_down(mx, o, 64));
  if ((threadIdx.x & 63) == 0) rmax[threadIdx.x >> 6] = mx;
  const double tot = block_sum(ss);
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) m = fmax(m, rmax[w]);
    res->resid_max = m;
    res->resid_rms = sqrt(tot / (double)nn);
  }
}

hipError_t resid_stats(const double* S, int n, StepDev* res, hipStream_t st) {
  hipLaunchKernelGGL(k_resid, dim3(1), dim3(1024), 0, st, S, n, res);
  return hipGetLastError();
}

__global__ void k_fill_zero(uint32_t* __restrict__ p, size_t n) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x)
    p[e] = 0u;
}

hipError_t fill_zero(void* p, size_t bytes, hipStream_t st) {
  const size_t n = bytes / 4;
  if (n == 0) return hipSuccess;
  int blocks = (int)((n + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, st, (uint32_t*)p, n);
  return hipGetLastError();
}

}  // namespace oa
