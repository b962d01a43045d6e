// oa_stockham.h -- Stockham autosort 1D FFTs on LDS for runtime lengths
// n = 2^a 3^b 4^c 5^d (radix 2/3/4/5 passes), batched over many transforms
// per workgroup.  Sign S = -1 is FFTW_FORWARD (exp(-2 pi i jk/n)), S = +1
// FFTW_BACKWARD; unnormalised.
#pragma once

#include "oa_common.h"

namespace oa {

// ---------------------------------------------------------------- butterflies
template <typename T, int R, int S>
struct Bfly;

template <typename T, int S>
struct Bfly<T, 2, S> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    cpx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <typename T, int S>
struct Bfly<T, 3, S> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    const T h = (T)0.86602540378443864676;  // sqrt(3)/2
    cpx<T> t1 = cadd(v[1], v[2]);
    cpx<T> t2 = {v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y};
    cpx<T> d = csub(v[1], v[2]);
    // i*S*h*d = (-S h d.y, S h d.x)
    cpx<T> r = {-(T)S * h * d.y, (T)S * h * d.x};
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, r);
    v[2] = csub(t2, r);
  }
};

template <typename T, int S>
struct Bfly<T, 4, S> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    cpx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    cpx<T> t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
    cpx<T> r = {-(T)S * t3.y, (T)S * t3.x};  // i*S*t3
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, r);
    v[3] = csub(t1, r);
  }
};

template <typename T, int S>
struct Bfly<T, 5, S> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
    cpx<T> b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
    cpx<T> d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
    cpx<T> a0 = v[0];
    cpx<T> p1 = {a0.x + c1 * b1.x + c2 * b2.x, a0.y + c1 * b1.y + c2 * b2.y};
    cpx<T> p2 = {a0.x + c2 * b1.x + c1 * b2.x, a0.y + c2 * b1.y + c1 * b2.y};
    cpx<T> q1 = {s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y};
    cpx<T> q2 = {s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y};
    cpx<T> iq1 = {-(T)S * q1.y, (T)S * q1.x};
    cpx<T> iq2 = {-(T)S * q2.y, (T)S * q2.x};
    v[0] = {a0.x + b1.x + b2.x, a0.y + b1.y + b2.y};
    v[1] = cadd(p1, iq1);
    v[4] = csub(p1, iq1);
    v[2] = cadd(p2, iq2);
    v[3] = csub(p2, iq2);
  }
};

// One Stockham pass: nbatch transforms of length n, stored at src[tr*ldb+e],
// result to dst.  tw[j] = exp(-2 pi i j / n).  NT: the workgroup size when
// known at compile time (0: blockDim.x, which is read from memory).
template <typename T, int R, int S, int NT = 0>
__device__ __forceinline__ void fft_pass(const cpx<T>* __restrict__ src, cpx<T>* __restrict__ dst,
                                         int n, int Ns, int nbatch, int ldb,
                                         const cpx<T>* __restrict__ tw) {
  const int m = n / R;
  const int total = m * nbatch;
  const int twstep = n / (Ns * R);
  const int step = NT > 0 ? NT : (int)blockDim.x;
  for (int g = threadIdx.x; g < total; g += step) {
    const int tr = g / m;
    const int j = g - tr * m;
    const int k = j % Ns;
    const cpx<T>* s = src + tr * ldb;
    cpx<T> v[R];
#pragma unroll
    for (int t = 0; t < R; t++) v[t] = s[j + t * m];
    if (Ns > 1) {
#pragma unroll
      for (int t = 1; t < R; t++) {
        cpx<T> w = tw[t * k * twstep];
        if (S > 0) w.y = -w.y;
        v[t] = cmul(v[t], w);
      }
    }
    Bfly<T, R, S>::run(v);
    const int idxD = (j - k) * R + k;
    cpx<T>* d = dst + tr * ldb;
#pragma unroll
    for (int t = 0; t < R; t++) d[idxD + t * Ns] = v[t];
  }
}

// Full 1D transform on LDS.  Caller has synchronised after filling `a`.
// Returns the buffer holding the result (a or b); synchronised on return.
template <typename T, int S, int NT = 0>
__device__ cpx<T>* fft_lds(cpx<T>* a, cpx<T>* b, const FftPlan& pl, int nbatch, int ldb,
                           const cpx<T>* tw) {
  // the plan's radices read once, all loads issued together (indexing the
  // plan by the pass number would cost one memory latency per pass)
  int rad[OA_MAXPASS];
#pragma unroll
  for (int q = 0; q < OA_MAXPASS; q++) rad[q] = pl.radix[q];
  const int npass = pl.npass;
  int Ns = 1;
  for (int p = 0; p < npass; p++) {
    int R = rad[0];
#pragma unroll
    for (int q = 1; q < OA_MAXPASS; q++) R = (p == q) ? rad[q] : R;
    switch (R) {
      case 2: fft_pass<T, 2, S, NT>(a, b, pl.n, Ns, nbatch, ldb, tw); break;
      case 3: fft_pass<T, 3, S, NT>(a, b, pl.n, Ns, nbatch, ldb, tw); break;
      case 4: fft_pass<T, 4, S, NT>(a, b, pl.n, Ns, nbatch, ldb, tw); break;
      default: fft_pass<T, 5, S, NT>(a, b, pl.n, Ns, nbatch, ldb, tw); break;
    }
    __syncthreads();
    cpx<T>* t = a; a = b; b = t;
    Ns *= R;
  }
  return a;
}

template <typename T, int NT = 0>
__device__ __forceinline__ void load_tw(cpx<T>* tw, const double2* __restrict__ g, int n) {
  const int step = NT > 0 ? NT : (int)blockDim.x;
  for (int i = threadIdx.x; i < n; i += step) tw[i] = {(T)g[i].x, (T)g[i].y};
}

}  // namespace oa
