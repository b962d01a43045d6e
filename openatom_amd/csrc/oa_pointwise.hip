// oa_pointwise.hip -- G-space pointwise kernels and small reductions of the
// CP electronic step.
//   k_sd       : psi += dt/m F  (CP_integrate_min_STD, cp_min_std.C:56-86)
//   k_cg_gamma, k_cg_update : conjugate-gradient direction and update
//                (CP_integrate_min_CG, cp_min_CG.C:55-105; reset rule
//                 CP_State_GSpacePlane.C:2646-2671)
//   k_red_sync : redundant kx=0 points <- conj(partner)
//                (CP_State_GSpacePlane.C:2862-3061)
//   k_reduce_to, k_dot, k_resid, k_fill_zero: deterministic reductions and
//   graph-safe zero fill.
#include <algorithm>

#include "oa_common.h"
#include "oa_kernels.h"

namespace oa {

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
                     const double* __restrict__ dtm, const int* __restrict__ twin, int nst,
                     int npts, int ld, double* __restrict__ fpart) {
  const size_t tot = (size_t)nst * npts;
  double fn = 0.0;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(e / npts), p = (int)(e - (size_t)s * npts);
    const size_t o = (size_t)s * ld + p;
    const double2 fv = f[o];
    const double m = dtm[p];
    const int tw = twin ? twin[p] : -1;
    double2 c = psi[o];
    c.x += fv.x * m;
    c.y += fv.y * m;
    if (tw != -2) psi[o] = c;  // a redundant point is written by its partner
    if (tw >= 0) psi[(size_t)s * ld + tw] = make_double2(c.x, -c.y);
    fn += fv.x * fv.x + fv.y * fv.y;
  }
  const double t = block_sum(fn);
  if (threadIdx.x == 0) fpart[blockIdx.x] = t;
}

hipError_t sd_update(double2* psi, const double2* f, const double* dtm, const int* twin, int nst,
                     int npts, int ld, double* fpart, int nfpart, hipStream_t st) {
  hipLaunchKernelGGL(k_sd, dim3(nfpart), dim3(256), 0, st, psi, f, dtm, twin, nst, npts, ld,
                     fpart);
  return hipGetLastError();
}

// acceptCgOverlap (CP_State_GSpacePlane.C:131-135) and the reset rule
// (:2651-2671): gamma = 0 on the first iteration, or when |F| grew by more
// than 5% after at least five steps without a reset, or more than doubled;
// else fovlap / fovlap_old.
__global__ void k_cg_gamma(const StepDev* __restrict__ res, CgDev* __restrict__ cg) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  CgDev s = *cg;
  s.fovlap_old = s.fovlap;
  s.fmag_old = s.fmag;
  s.fovlap = res->fnorm2;
  s.fmag = sqrt(s.fovlap);
  s.iter++;
  int reset = 0;
  if (s.iter == 1) {
    reset = 1;
    s.num_reset = 0;
  } else if ((s.fmag > 1.05 * s.fmag_old && s.num_reset >= 5) || s.fmag > 2.0 * s.fmag_old) {
    reset = 1;
    s.num_reset = 0;
  }
  s.num_reset++;
  s.gamma = reset ? 0.0 : s.fovlap / s.fovlap_old;
  *cg = s;
}

hipError_t cg_gamma(const StepDev* res, CgDev* cg, hipStream_t st) {
  hipLaunchKernelGGL(k_cg_gamma, dim3(1), dim3(64), 0, st, res, cg);
  return hipGetLastError();
}

// cg = gamma cg + F where g2 <= ecut; psi += dt/m cg (cp_min_CG.C:55-105)
__global__ void k_cg_update(double2* __restrict__ psi, double2* __restrict__ cgd,
                            const double2* __restrict__ f, const double* __restrict__ dtm,
                            const double* __restrict__ g2, double ecut, int nst, int npts, int ld,
                            const CgDev* __restrict__ cg) {
  const double gamma = cg->gamma;
  const size_t tot = (size_t)nst * npts;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(e / npts), p = (int)(e - (size_t)s * npts);
    const size_t o = (size_t)s * ld + p;
    double2 d = cgd[o];
    if (g2[p] <= ecut) {
      const double2 fv = f[o];
      d.x = d.x * gamma + fv.x;
      d.y = d.y * gamma + fv.y;
      cgd[o] = d;
    }
    const double m = dtm[p];
    double2 c = psi[o];
    c.x += d.x * m;
    c.y += d.y * m;
    psi[o] = c;
  }
}

hipError_t cg_update(double2* psi, double2* cgdir, const double2* f, const double* dtm,
                     const double* g2, double ecut, int nst, int npts, int ld, const CgDev* cg,
                     hipStream_t st) {
  const size_t tot = (size_t)nst * npts;
  int blocks = (int)std::min<size_t>((tot + 255) / 256, 4096);
  hipLaunchKernelGGL(k_cg_update, dim3(blocks), dim3(256), 0, st, psi, cgdir, f, dtm, g2, ecut,
                     nst, npts, ld, cg);
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

// One workgroup; each thread's loads are issued eight at a time
// (branch-free: index clamped, value masked) rather than one per iteration
constexpr int kRedThreads = 1024;
__global__ __launch_bounds__(kRedThreads) void k_reduce_to(const double* __restrict__ part, int n,
                                                           double scale,
                                                           double* __restrict__ out) {
  double v = 0.0;
  for (int i0 = 0; i0 < n; i0 += 8 * kRedThreads) {
    double a[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int i = i0 + (int)threadIdx.x + u * kRedThreads;
      const double w = part[min(i, n - 1)];
      a[u] = i < n ? w : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) v += a[u];
  }
  const double t = block_sum(v);
  if (threadIdx.x == 0) *out = t * scale;
}

hipError_t reduce_to(const double* part, int n, double scale, double* out, hipStream_t st) {
  if (n < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_reduce_to, dim3(1), dim3(kRedThreads), 0, st, part, n, scale, out);
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
