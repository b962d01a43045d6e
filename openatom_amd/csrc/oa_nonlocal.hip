// oa_nonlocal.hip -- non-local Kleinman-Bylander pseudopotential, l = 0
// channel, by the reference's G-space method (its default; the EES
// real-space method is not built).  The structure-factor table is built once
// per atom configuration; the per-step projection Z = Psi SF^T and the
// force F -= w Z' SF run on the FP64 matrix-core gram and rotate kernels
// (oa_gemm.hip).  This file holds the table builder and the two small
// epilogues: energy + scaled transpose, and the atom forces.
#include <hip/hip_runtime.h>

#include "oa_kernels.h"

namespace oa {
namespace {

// get_rad_proj (cp_struct_fact.C:268-300): cubic spline row
// i = clamp(floor((g - gmin)/dg) + 1, 1, nspline), h = g - ((i-1) dg + gmin);
// the G = 0 value is the table's own
__device__ __forceinline__ double nl_proj(const NlDev& d, int sp, double g) {
  if (g == 0.0) return d.proj_g0[sp];
  const int ns = d.nspline;
  int i = (int)((g - d.gmin) / d.dg) + 1;
  i = min(max(i, 1), ns);
  const double h = g - ((double)(i - 1) * d.dg + d.gmin);
  const double* c = d.proj + (size_t)sp * 4 * ns + (i - 1);
  return ((c[3 * ns] * h + c[2 * ns]) * h + c[ns]) * h + c[0];
}

// One thread per stored point, atoms [blockIdx.y * kNlAtoms, +kNlAtoms).
// Rows q * natoms + I of sf (interleaved re, im):
//   q = 0: P y00 exp(-i G.R)        (the conjugated table of
//          CP_calc_Struct_Fact, cp_struct_fact.C:186-256, without w)
//   q = 1..3: conj(i G_q P y00 exp(i G.R)) = -G_q P y00 (sin, cos)(G.R),
//          so that Re(psi conj(row)) = d/dR_q Re(psi P y00 exp(i G.R))
constexpr int kNlAtoms = 8;
__global__ __launch_bounds__(256) void k_nl_sf(NlDev d, double* __restrict__ sf) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= d.npts) return;
  const double tpi = 2.0 * M_PI;
  const int a = d.kx[p], b = d.ky[p], c = d.kz[p];
  const double* hi = d.hmati;
  const double gx = tpi * (a * hi[0] + b * hi[1] + c * hi[2]);
  const double gy = tpi * (a * hi[3] + b * hi[4] + c * hi[5]);
  const double gz = tpi * (a * hi[6] + b * hi[7] + c * hi[8]);
  const double g = sqrt(gx * gx + gy * gy + gz * gz);
  const int i0 = blockIdx.y * kNlAtoms, i1 = min(d.natoms, i0 + kNlAtoms);
  const size_t rs = (size_t)d.natoms * d.ldr;  // between the four row groups
  for (int ia = i0; ia < i1; ia++) {
    const double* R = d.pos + 3 * ia;
    const double arg = R[0] * gx + R[1] * gy + R[2] * gz;
    double sn, cs;
    sincos(arg, &sn, &cs);
    const double pr = nl_proj(d, d.species[ia], g) * d.y00;
    double* row = sf + (size_t)ia * d.ldr + 2 * (size_t)p;
    row[0] = pr * cs;
    row[1] = -pr * sn;
    row[rs] = -gx * pr * sn;
    row[rs + 1] = -gx * pr * cs;
    row[2 * rs] = -gy * pr * sn;
    row[2 * rs + 1] = -gy * pr * cs;
    row[3 * rs] = -gz * pr * sn;
    row[3 * rs + 1] = -gz * pr * cs;
  }
}

// CP_enl_atm_forc_calc (cp_nl_energy_forc.C:196-216) for all states:
// Mt[I][s] = 2 vn[I] Z[s][I]; *enl = sum_s occ[s] sum_I vn[I] Z[s][I]^2
// (vn = vnorm/Omega), one workgroup, fixed summation order
__global__ __launch_bounds__(256) void k_nl_energy(const double* __restrict__ Z, int ldz, int n,
                                                   int natm, const double* __restrict__ vn,
                                                   const double* __restrict__ occ,
                                                   double* __restrict__ Mt, double* enl) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  const int tot = n * natm;
  double e = 0.0;
  for (int k = tid; k < tot; k += 256) {
    const int s = k / natm, ia = k - s * natm;
    const double z = Z[(size_t)s * ldz + ia];
    e += occ[s] * vn[ia] * z * z;
    Mt[(size_t)ia * n + s] = 2.0 * vn[ia] * z;
  }
  red[tid] = e;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) *enl = red[0];
}

// f[I][q] = -sum_s 2 vn[I] Z[s][I] dZ_q[s][I] (cp_nl_energy_forc.C:219-229);
// Z: [n][4 natm] with the derivative columns after the projections
__global__ __launch_bounds__(64) void k_nl_atom_forces(const double* __restrict__ Z, int n,
                                                       int natm, const double* __restrict__ vn,
                                                       double* __restrict__ f) {
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (t >= 3 * natm) return;
  const int ia = t / 3, q = t - 3 * ia;
  const int ld = 4 * natm;
  double acc = 0.0;
  for (int s = 0; s < n; s++)
    acc += (2.0 * vn[ia] * Z[(size_t)s * ld + ia]) * Z[(size_t)s * ld + (q + 1) * natm + ia];
  f[t] = -acc;
}

}  // namespace

hipError_t nl_sf_build(const NlDev& d, double* sf, hipStream_t st) {
  if (d.natoms < 1 || d.npts < 1 || d.ldr < 2 * d.npts) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_nl_sf, dim3((d.npts + 255) / 256, (d.natoms + kNlAtoms - 1) / kNlAtoms),
                     dim3(256), 0, st, d, sf);
  return hipGetLastError();
}

hipError_t nl_energy(const double* Z, int ldz, int n, int natm, const double* vn,
                     const double* occ, double* Mt, double* enl, hipStream_t st) {
  hipLaunchKernelGGL(k_nl_energy, dim3(1), dim3(256), 0, st, Z, ldz, n, natm, vn, occ, Mt, enl);
  return hipGetLastError();
}

hipError_t nl_atom_forces(const double* Z, int n, int natm, const double* vn, double* f,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_nl_atom_forces, dim3((3 * natm + 63) / 64), dim3(64), 0, st, Z, n, natm,
                     vn, f);
  return hipGetLastError();
}

}  // namespace oa
