// oa_gemm.hip -- dense FP64 contractions of the CP electronic step on
// CDNA4 matrix cores (v_mfma_f64_16x16x4f64).
//
//   gram    : S = 2 Re Psi^H Psi over stored points with kx=0 halved
//             (PairCalculator::multiplyForward, ckPairCalculator.C:682-802)
//             and Lambda = Re Psi^H F (asymmetric PairCalculator, alpha = 1);
//             split-K over the long point axis, the symmetric case computes
//             the upper tiles only.
//   rotate  : Psi' = Psi T and F -= w Psi Lambda
//             (bwMultiplyHelper, ckPairCalculator.C:1189-1425;
//              unpackLambda, CP_State_GSpacePlane.C:2045-2225).
//   ns      : Newton-Schulz iteration for (S/2)^{-1/2}
//             (Ortho::do_iteration/step_2/step_3/tolerance_check,
//              ortho.C:862-985; serial twin cp_lowdin.C:180-265), either one
//             cooperative launch that loops on the device with grid barriers
//             or, as a fallback, two launches per iteration with a device-side
//             early exit.
//
// All three feed the matrix cores straight from global memory (no LDS
// staging): inside one MFMA step the assignment of the four k values to the
// four lane groups is free as long as A and B agree, so lane group g takes
// the contiguous run k0+4g..k0+4g+3 and each lane issues 32-byte loads.
#include <hip/hip_cooperative_groups.h>

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
// part[sp][i][j] = sum_{k in slice sp} w(k) A[i][k] B[j][k].  One workgroup
// = one 32x32 output tile and one K slice; its four waves take a quarter of
// the slice each and are summed through LDS.
#define GT 32

__global__ __launch_bounds__(256) void k_gram(GramArgs a, int ntile) {
  __shared__ double red[4][GT][GT + 1];
  int ti, tj;
  if (a.sym) {
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
  const int l16 = lane & 15, g = lane >> 4;
  const double* B = a.sym ? a.A : a.B;
  const int kbeg = sp * a.kslice;
  const int kend = min(a.K, kbeg + a.kslice);
  const int q = ((a.kslice / 4) + 15) & ~15;
  const int wb = kbeg + wave * q, we = min(kend, wb + q);
  int ra[2], rb[2];
  bool va[2], vb[2];
#pragma unroll
  for (int x = 0; x < 2; x++) {
    ra[x] = ti * GT + x * 16 + l16;
    rb[x] = tj * GT + x * 16 + l16;
    va[x] = ra[x] < a.n;
    vb[x] = rb[x] < a.n;
  }
  v4d acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) acc[x][y] = (v4d){0.0, 0.0, 0.0, 0.0};
  for (int k0 = wb; k0 < we; k0 += 16) {
    const int kk = k0 + 4 * g;
    double av[2][4], bv[2][4];
    if (k0 + 16 <= we) {
#pragma unroll
      for (int x = 0; x < 2; x++) {
        const double* pa = a.A + (size_t)(va[x] ? ra[x] : 0) * a.lda + kk;
        const double* pb = B + (size_t)(vb[x] ? rb[x] : 0) * a.lda + kk;
#pragma unroll
        for (int e = 0; e < 4; e++) {
          av[x][e] = va[x] ? pa[e] : 0.0;
          bv[x][e] = vb[x] ? pb[e] : 0.0;
        }
      }
    } else {
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int k = kk + e;
          av[x][e] = (va[x] && k < we) ? a.A[(size_t)ra[x] * a.lda + k] : 0.0;
          bv[x][e] = (vb[x] && k < we) ? B[(size_t)rb[x] * a.lda + k] : 0.0;
        }
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const double w = (kk + e < a.k_lo) ? a.w_lo : a.w_hi;
#pragma unroll
      for (int x = 0; x < 2; x++) av[x][e] *= w;
    }
#pragma unroll
    for (int e = 0; e < 4; e++)
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) acc[x][y] = mfma16(av[x][e], bv[y][e], acc[x][y]);
  }
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) red[wave][x * 16 + g + 4 * v][y * 16 + l16] = acc[x][y][v];
  __syncthreads();
  for (int e = tid; e < GT * GT; e += 256) {
    const int r = e / GT, c = e - r * GT;
    const int row = ti * GT + r, col = tj * GT + c;
    if (row < a.n && col < a.n)
      a.part[((size_t)sp * a.n + row) * a.n + col] =
          red[0][r][c] + red[1][r][c] + red[2][r][c] + red[3][r][c];
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
    const double* p = part + (size_t)si * n + sj;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int sp = 0;
    for (; sp + 4 <= nsplit; sp += 4) {
      a0 += p[(size_t)sp * nn];
      a1 += p[(size_t)(sp + 1) * nn];
      a2 += p[(size_t)(sp + 2) * nn];
      a3 += p[(size_t)(sp + 3) * nn];
    }
    for (; sp < nsplit; sp++) a0 += p[(size_t)sp * nn];
    C[e] = alpha * ((a0 + a1) + (a2 + a3));
  }
}

hipError_t gram_reduce(const double* part, int nsplit, int n, int sym, double alpha, double* C,
                       hipStream_t st) {
  const size_t nn = (size_t)n * n;
  int blocks = (int)((nn + 255) / 256);
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_gram_reduce, dim3(blocks), dim3(256), 0, st, part, nsplit, n, sym, alpha,
                     C);
  return hipGetLastError();
}

// ------------------------------------------------------------------ rotate
// Y[j][k] = beta Y[j][k] + s(k) sum_i M[i][j] X[i][k].  Workgroup: 32 columns
// of k and up to 128 rows of j; wave w owns j rows [32w, 32w+32).
__global__ __launch_bounds__(256) void k_rotate(RotArgs a) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const int k0 = blockIdx.x * 32;
  const int jw = blockIdx.y * 128 + wave * 32;
  if (jw >= a.n) return;
  v4d acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) acc[x][y] = (v4d){0.0, 0.0, 0.0, 0.0};
  const int kc[2] = {k0 + l16, k0 + 16 + l16};
  const int jc[2] = {jw + l16, jw + 16 + l16};
  for (int i0 = 0; i0 < a.n; i0 += 4) {
    const int i = i0 + g;
    const bool vi = i < a.n;
    double am[2], bx[2];
#pragma unroll
    for (int x = 0; x < 2; x++) am[x] = (vi && jc[x] < a.n) ? a.M[(size_t)i * a.n + jc[x]] : 0.0;
#pragma unroll
    for (int y = 0; y < 2; y++) bx[y] = (vi && kc[y] < a.K) ? a.X[(size_t)i * a.ldx + kc[y]] : 0.0;
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 2; y++) acc[x][y] = mfma16(am[x], bx[y], acc[x][y]);
  }
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int j = jw + x * 16 + g + 4 * v;
        const int k = k0 + y * 16 + l16;
        if (j < a.n && k < a.K) {
          const double s = (k < a.k_lo) ? a.s_lo : a.s_hi;
          double* yp = a.Y + (size_t)j * a.ldx + k;
          const double val = s * acc[x][y][v];
          *yp = (a.beta != 0.0) ? a.beta * (*yp) + val : val;
        }
      }
}

hipError_t rotate_f64(const RotArgs& a, hipStream_t st) {
  dim3 grid((a.K + 31) / 32, (a.n + 127) / 128);
  hipLaunchKernelGGL(k_rotate, grid, dim3(256), 0, st, a);
  return hipGetLastError();
}

// ------------------------------------------------------------------ Newton-Schulz
// Workgroup = 32x32 tile of the n x n matrices, wave = 16x16 block.
#define NT 32

// 16x16 block (row0, col0) of A B, A and B n x n row-major
__device__ __forceinline__ v4d blk_gemm(const double* __restrict__ A,
                                        const double* __restrict__ B, int n, int row0,
                                        int col0) {
  const int lane = threadIdx.x & 63, l16 = lane & 15, g = lane >> 4;
  const int r = row0 + l16, cc = col0 + l16;
  v4d acc = (v4d){0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < n; k0 += 4) {
    const int k = k0 + g;
    const double av = (r < n && k < n) ? A[(size_t)r * n + k] : 0.0;
    const double bv = (k < n && cc < n) ? B[(size_t)k * n + cc] : 0.0;
    acc = mfma16(av, bv, acc);
  }
  return acc;
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
// X0 = S/2, T0 = I; loop { C = 3I - T X; X' = C X ... } with two grid
// barriers per iteration; stops when the rms change of T <= tol.
__global__ __launch_bounds__(256) void k_ns_coop(NsArgs a) {
  namespace cg = cooperative_groups;
  cg::grid_group grid = cg::this_grid();
  const int n = a.n, nt = (n + NT - 1) / NT, nwg = nt * nt;
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
    blk_store(a.C, blk_gemm(T[o], X[o], n, row0, col0), n, row0, col0, -1.0, 3.0, nullptr);
    grid.sync();
    blk_store(X[w], blk_gemm(X[o], a.C, n, row0, col0), n, row0, col0, 0.5, 0.0, nullptr);
    const double e =
        blk_store(T[w], blk_gemm(a.C, T[o], n, row0, col0), n, row0, col0, 0.5, 0.0, T[o]);
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

// ---- fallback: two launches per iteration with a device-side early exit

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

// True when iteration `it` must not run (converged before it).  Thread 0
// decides and broadcasts, so the whole workgroup takes the same path.
__device__ bool ns_skip(const double* __restrict__ errpart, int it, int nwg, int n, double tol,
                        StepDev* res, bool record) {
  __shared__ int skip;
  if (threadIdx.x == 0) {
    int f = 0;
    if (*(volatile int*)&res->ns_iter != 0) {
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
                                                  StepDev* res) {
  const int nt = (n + NT - 1) / NT, nwg = nt * nt;
  if (ns_skip(errpart, it, nwg, n, tol, res, true)) return;
  int row0, col0;
  blk_coords(nt, row0, col0);
  blk_store(C, blk_gemm(T, X, n, row0, col0), n, row0, col0, -1.0, 3.0, nullptr);
}

__global__ __launch_bounds__(256) void k_ns_step23(const double* __restrict__ Xo,
                                                   const double* __restrict__ To,
                                                   const double* __restrict__ C,
                                                   double* __restrict__ Xn,
                                                   double* __restrict__ Tn, int n, int it,
                                                   double tol, double* errpart, StepDev* res) {
  const int nt = (n + NT - 1) / NT, nwg = nt * nt;
  if (ns_skip(errpart, it, nwg, n, tol, res, false)) return;
  int row0, col0;
  blk_coords(nt, row0, col0);
  blk_store(Xn, blk_gemm(Xo, C, n, row0, col0), n, row0, col0, 0.5, 0.0, nullptr);
  const double e = blk_store(Tn, blk_gemm(C, To, n, row0, col0), n, row0, col0, 0.5, 0.0, To);
  const double es = wg_sum(e);
  if (threadIdx.x == 0) errpart[(size_t)it * nwg + blockIdx.x] = es;
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

static hipError_t ns_run_launches(const NsArgs& a, hipStream_t st) {
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
    hipLaunchKernelGGL(k_ns_step1, dim3(nwg), dim3(256), 0, st, T[o], X[o], a.C, n, it, a.tol,
                       a.errpart, a.res);
    hipLaunchKernelGGL(k_ns_step23, dim3(nwg), dim3(256), 0, st, X[o], T[o], a.C, X[w], T[w], n,
                       it, a.tol, a.errpart, a.res);
  }
  hipLaunchKernelGGL(k_ns_final, dim3(1), dim3(64), 0, st, a.errpart, nwg, n, a.max_iter, a.tol,
                     a.res);
  hipLaunchKernelGGL(k_ns_copy, dim3(eb), dim3(256), 0, st, a.T0, a.T1, a.Tout, n, a.res);
  return hipGetLastError();
}

// Can all n x n tiles be resident at once for the cooperative launch?
bool ns_coop_possible(int n, int device) {
  const int nt = (n + NT - 1) / NT;
  int per_cu = 0, ncu = 0, coop = 0;
  if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device) != hipSuccess ||
      !coop)
    return false;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return false;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_ns_coop, 256, 0) !=
      hipSuccess)
    return false;
  return (long)per_cu * ncu >= (long)nt * nt;
}

hipError_t ns_run(const NsArgs& a, hipStream_t st) {
  if (!a.coop) return ns_run_launches(a, st);
  const int nt = (a.n + NT - 1) / NT;
  NsArgs args = a;
  void* params[] = {&args};
  return hipLaunchCooperativeKernel((const void*)k_ns_coop, dim3(nt * nt), dim3(256), params, 0,
                                    st);
}

}  // namespace oa
