// oa_gemm_common.h -- helpers shared by the FP64 (oa_gemm.hip) and the
// split-BF16 (oa_gemm_bf16.hip) contractions: compile-time loops, the K
// slices of the gram products and the register ring of operand chunks.
#pragma once

#include <algorithm>
#include <type_traits>

#include "oa_kernels.h"

namespace oa {

// f(integral_constant<int, I>) for I in [B, E): compile-time indices into
// register arrays
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// One K slice: chunks [c0, c1) of 16 k values starting at `base`; loads are
// masked to the region [rs, re), which is either the kx=0 part [0, k_lo) or
// the rest [k_lo, K), so a slice has one weight.
struct GSlice {
  int rs, re, base, c0, c1;
  double w;
};

__device__ __forceinline__ GSlice gram_slice(const GramArgs& a, int sp, int nlo) {
  GSlice s;
  int idx, cnt;
  const int klo = min(max(a.k_lo, 0), a.K);
  if (sp < nlo) {
    s.rs = 0;
    s.re = klo;
    idx = sp;
    cnt = nlo;
    s.w = a.w_lo;
  } else {
    s.rs = klo;
    s.re = a.K;
    idx = sp - nlo;
    cnt = a.nslice - nlo;
    s.w = a.w_hi;
  }
  s.base = s.rs & ~15;
  const int nch = (s.re - s.base + 15) >> 4;
  s.c0 = (int)(((long long)idx * nch) / cnt);
  s.c1 = (int)(((long long)(idx + 1) * nch) / cnt);
  return s;
}

// Register ring of D chunk buffers: the loads of chunk c + D are issued as
// soon as the MFMAs of chunk c have been issued, so D - 1 chunks of MFMAs
// cover the memory latency.  load(c, buf) fills a buffer for chunk c;
// mma(c, buf) issues the chunk's MFMAs.  Loads are issued unconditionally
// (past the slice they repeat its last chunk): a load that only some paths
// issue forces the compiler to wait for every outstanding load.
template <int D, typename Buf, typename L, typename M>
__device__ __forceinline__ void chunk_ring(int c0, int c1, Buf (&ring)[D], L&& load, M&& mma) {
  if (c0 >= c1) return;
  const int cl = c1 - 1;
  static_for<0, D>([&](auto dc) __attribute__((always_inline)) {
    constexpr int d = decltype(dc)::value;
    load(min(c0 + d, cl), ring[d]);
  });
  for (int c = c0; c < c1; c += D) {
    static_for<0, D>([&](auto dc) __attribute__((always_inline)) {
      constexpr int d = decltype(dc)::value;
      if (c + d < c1) mma(c + d, ring[d]);
      load(min(c + d + D, cl), ring[d]);
    });
  }
}

// K slices of the kx=0 region (weight w_lo) among nslice
inline int gram_nlo(const GramArgs& a) {
  if (a.k_lo <= 0) return 0;
  if (a.k_lo >= a.K) return a.nslice;
  const int nlo = (int)((double)a.nslice * a.k_lo / a.K + 0.5);
  return std::min(std::max(nlo, 1), a.nslice - 1);
}

}  // namespace oa
