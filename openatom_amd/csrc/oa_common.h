// oa_common.h -- shared device/host definitions for the CP electronic-step
// library (CDNA4 / gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define OA_NTHR 256      // threads per workgroup for the FFT and pointwise kernels
#define OA_MAXPASS 8     // max Stockham passes per 1D transform
#define OA_XSLOTS 8      // register slots per thread in the x passes: pair-rows * nx <= OA_XSLOTS * OA_NTHR

template <typename T>
struct cpx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {  // a * conj(b)
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cconj(cpx<T> a) { return {a.x, -a.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return {a.x * s, a.y * s}; }

// One 1D transform length decomposed into Stockham radix passes.
struct FftPlan {
  int n;
  int npass;
  int radix[OA_MAXPASS];
};

// Per-system geometry shared by the FFT kernels (passed by value).
struct Geom {
  int nx, ny, nz;     // FFT grid
  int nplane_x;       // kx columns carried in the half plane (kx_max + 1)
  int nlines;         // z-lines of the sphere
  int npts;           // stored points per state
  int ld;             // complex stride between states in G-space arrays (>= npts)
  int nst;            // states
};
