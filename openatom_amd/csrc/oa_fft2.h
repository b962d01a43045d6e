// oa_fft2.h -- register-resident two-factor FFT engine for the state FFTs.
//
// A length N = R1*R2 transform is done by a group of max(R1, R2) threads:
//   pass 1: thread j (< R2) holds x[j + t*R2] (t < R1) in registers and runs
//           a radix-R1 DFT; its outputs land at j*R1 + t of an LDS buffer;
//   pass 2: thread j' (< R1) reads buf[j' + t*R1] (t < R2), multiplies by
//           W_N^{t j'} and runs a radix-R2 DFT; X[j' + t*R1] is the result
//           in natural order (the two Stockham passes with Ns = 1, R1).
// One LDS exchange per transform.  The thread -> (transform, j) map of each
// pass is chosen per kernel so that the global loads and stores of that pass
// are contiguous across lanes.  Radix-R DFTs up to 16 are unrolled in
// registers: 2, 3, 4, 5 directly, composites of coprime factors by the
// prime-factor map (no twiddles), the others as P x Q Cooley-Tukey with
// compile-time twiddles.
#pragma once

#include "oa_common.h"

namespace oa2 {

using ::cpx;

// ---------------------------------------------------------------- constants
constexpr double kTwoPi = 6.283185307179586476925286766559;

constexpr double ct_sin_small(double x) {  // |x| <= pi/4
  double x2 = x * x, term = x, sum = x;
  for (int k = 1; k < 14; ++k) {
    term *= -x2 / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}

constexpr double ct_cos_small(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; ++k) {
    term *= -x2 / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}

// cos and sin of 2 pi m / n, reduced exactly to |delta| <= pi/4
constexpr void ct_cs(int m, int n, double& c, double& s) {
  m %= n;
  if (m < 0) m += n;
  // nearest quarter turn q: 4m/n rounded
  int q = (8 * m + n) / (2 * n);
  const double delta = kTwoPi * (double)(4 * m - q * n) / (4.0 * n);
  const double cd = ct_cos_small(delta), sd = ct_sin_small(delta);
  switch (q & 3) {
    case 0: c = cd; s = sd; break;
    case 1: c = -sd; s = cd; break;
    case 2: c = -cd; s = -sd; break;
    default: c = sd; s = -cd; break;
  }
}

template <int R>
struct TwTab {
  double c[R];
  double s[R];
};

template <int R>
constexpr TwTab<R> make_tw() {
  TwTab<R> t{};
  for (int m = 0; m < R; ++m) ct_cs(m, R, t.c[m], t.s[m]);
  return t;
}

// ---------------------------------------------------------------- radix DFTs
template <typename T>
__device__ __forceinline__ cpx<T> cm(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

template <typename T, int R, int S>
struct Dft;

template <typename T, int S>
struct Dft<T, 1, S> {
  static __device__ __forceinline__ void run(cpx<T>*) {}
};

template <typename T, int S>
struct Dft<T, 2, S> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    const cpx<T> a = v[0], b = v[1];
    v[0] = {a.x + b.x, a.y + b.y};
    v[1] = {a.x - b.x, a.y - b.y};
  }
};

template <typename T, int S>
struct Dft<T, 3, S> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    const T h = (T)0.86602540378443864676;
    const cpx<T> t1 = {v[1].x + v[2].x, v[1].y + v[2].y};
    const cpx<T> t2 = {v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y};
    const cpx<T> d = {v[1].x - v[2].x, v[1].y - v[2].y};
    const cpx<T> r = {-(T)S * h * d.y, (T)S * h * d.x};
    v[0] = {v[0].x + t1.x, v[0].y + t1.y};
    v[1] = {t2.x + r.x, t2.y + r.y};
    v[2] = {t2.x - r.x, t2.y - r.y};
  }
};

template <typename T, int S>
struct Dft<T, 4, S> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    const cpx<T> t0 = {v[0].x + v[2].x, v[0].y + v[2].y}, t1 = {v[0].x - v[2].x, v[0].y - v[2].y};
    const cpx<T> t2 = {v[1].x + v[3].x, v[1].y + v[3].y}, t3 = {v[1].x - v[3].x, v[1].y - v[3].y};
    const cpx<T> r = {-(T)S * t3.y, (T)S * t3.x};
    v[0] = {t0.x + t2.x, t0.y + t2.y};
    v[2] = {t0.x - t2.x, t0.y - t2.y};
    v[1] = {t1.x + r.x, t1.y + r.y};
    v[3] = {t1.x - r.x, t1.y - r.y};
  }
};

template <typename T, int S>
struct Dft<T, 5, S> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
    const cpx<T> b1 = {v[1].x + v[4].x, v[1].y + v[4].y}, b2 = {v[2].x + v[3].x, v[2].y + v[3].y};
    const cpx<T> d1 = {v[1].x - v[4].x, v[1].y - v[4].y}, d2 = {v[2].x - v[3].x, v[2].y - v[3].y};
    const cpx<T> a0 = v[0];
    const cpx<T> p1 = {a0.x + c1 * b1.x + c2 * b2.x, a0.y + c1 * b1.y + c2 * b2.y};
    const cpx<T> p2 = {a0.x + c2 * b1.x + c1 * b2.x, a0.y + c2 * b1.y + c1 * b2.y};
    const cpx<T> q1 = {s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y};
    const cpx<T> q2 = {s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y};
    const cpx<T> iq1 = {-(T)S * q1.y, (T)S * q1.x}, iq2 = {-(T)S * q2.y, (T)S * q2.x};
    v[0] = {a0.x + b1.x + b2.x, a0.y + b1.y + b2.y};
    v[1] = {p1.x + iq1.x, p1.y + iq1.y};
    v[4] = {p1.x - iq1.x, p1.y - iq1.y};
    v[2] = {p2.x + iq2.x, p2.y + iq2.y};
    v[3] = {p2.x - iq2.x, p2.y - iq2.y};
  }
};

// R = P*Q: n = Q n1 + n2, k = k1 + P k2;
// X[k1 + P k2] = sum_n2 W_Q^{n2 k2} W_R^{n2 k1} sum_n1 x[Q n1 + n2] W_P^{n1 k1}
template <typename T, int P, int Q, int S>
struct DftPQ {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    constexpr int R = P * Q;
    constexpr TwTab<R> tw = make_tw<R>();
    cpx<T> y[Q][P];
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) {
      cpx<T> a[P];
#pragma unroll
      for (int n1 = 0; n1 < P; ++n1) a[n1] = v[Q * n1 + n2];
      Dft<T, P, S>::run(a);
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) {
        const int m = (n2 * k1) % R;
        if (m == 0) {
          y[n2][k1] = a[k1];
        } else {
          const cpx<T> w = {(T)tw.c[m], (T)(S * tw.s[m])};
          y[n2][k1] = cm(a[k1], w);
        }
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
      cpx<T> b[Q];
#pragma unroll
      for (int n2 = 0; n2 < Q; ++n2) b[n2] = y[n2][k1];
      Dft<T, Q, S>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) v[k1 + P * k2] = b[k2];
    }
  }
};

// R = P*Q with gcd(P, Q) = 1 (Good-Thomas prime-factor map, no twiddles):
// n = (Q n1 + P n2) mod R, k = (Q a k1 + P b k2) mod R with a = Q^-1 mod P,
// b = P^-1 mod Q, so that W_R^{nk} = W_P^{n1 k1} W_Q^{n2 k2}.
constexpr int ct_inv_mod(int x, int m) {
  for (int i = 1; i < m; ++i)
    if ((x * i) % m == 1) return i;
  return 1;
}

template <typename T, int P, int Q, int S>
struct DftPFA {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    constexpr int R = P * Q;
    constexpr int A = ct_inv_mod(Q % P, P), B = ct_inv_mod(P % Q, Q);
    cpx<T> y[Q][P];
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) {
      cpx<T> a[P];
#pragma unroll
      for (int n1 = 0; n1 < P; ++n1) a[n1] = v[(Q * n1 + P * n2) % R];
      Dft<T, P, S>::run(a);
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) y[n2][k1] = a[k1];
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
      cpx<T> b[Q];
#pragma unroll
      for (int n2 = 0; n2 < Q; ++n2) b[n2] = y[n2][k1];
      Dft<T, Q, S>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) v[(Q * A * k1 + P * B * k2) % R] = b[k2];
    }
  }
};

template <typename T, int S> struct Dft<T, 6, S> : DftPFA<T, 2, 3, S> {};
template <typename T, int S> struct Dft<T, 8, S> : DftPQ<T, 2, 4, S> {};
template <typename T, int S> struct Dft<T, 9, S> : DftPQ<T, 3, 3, S> {};
template <typename T, int S> struct Dft<T, 10, S> : DftPFA<T, 2, 5, S> {};
template <typename T, int S> struct Dft<T, 12, S> : DftPFA<T, 4, 3, S> {};
template <typename T, int S> struct Dft<T, 15, S> : DftPFA<T, 3, 5, S> {};
template <typename T, int S> struct Dft<T, 16, S> : DftPQ<T, 4, 4, S> {};

// pass-2 twiddle W_N^{m} from the table exp(-2 pi i j / N), conjugated for S = +1
template <typename T, int S>
__device__ __forceinline__ cpx<T> twiddle(const double2* __restrict__ tw, int m) {
  const double2 w = tw[m];
  return {(T)w.x, (T)(S < 0 ? w.y : -w.y)};
}

// Pruning to the sphere.  A grid that holds the density sphere |G|^2 <= 4
// Ecut has N >= 2 kr + 1 with kr >= 2 kmax, so the wavefunction components
// satisfy |k| <= KM = (N - 1) / 4 along every axis.  A radix pass reading
// (or writing) index i = j + t*STRIDE, j < STRIDE, touches only zeros (or
// unneeded outputs) for every j when KM < t*STRIDE and t*STRIDE + STRIDE - 1
// < N - KM, i.e. for t in [T0, T1].  Pruned kernels set those inputs to zero
// at compile time and do not store those outputs, so the compiler drops
// their arithmetic (the engine's translation units are built with
// -fno-signed-zeros, which lets x + 0 fold to x; only the sign of zero
// results can differ).
template <int N, int STRIDE>
struct Sph {
  static constexpr int KM = (N - 1) / 4;
  static constexpr int T0 = KM / STRIDE + 1;
  static constexpr int T1 = (N - KM - STRIDE) / STRIDE;
  static __device__ __forceinline__ constexpr bool cut(bool on, int t) {
    return on && t >= T0 && t <= T1;
  }
};

}  // namespace oa2
