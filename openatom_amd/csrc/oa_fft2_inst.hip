// One transform length of the two-factor FFT engine per translation unit:
// built with -DOA_FFT2_N=<N> -DOA_FFT2_R1=<R1> -DOA_FFT2_R2=<R2>.
#include "oa_fft2_kernels.h"

#if !defined(OA_FFT2_N) || !defined(OA_FFT2_R1) || !defined(OA_FFT2_R2)
#error "OA_FFT2_N, OA_FFT2_R1 and OA_FFT2_R2 must be defined"
#endif
static_assert(OA_FFT2_R1 * OA_FFT2_R2 == OA_FFT2_N, "N must be R1 * R2");

#define OA_CAT2(a, b) a##b
#define OA_CAT(a, b) OA_CAT2(a, b)

namespace oa2 {
hipError_t OA_CAT(launch_n, OA_FFT2_N)(int kind, bool fp32, const oa::FftCtx& c, hipStream_t st) {
  return fp32 ? launch<float, OA_FFT2_R1, OA_FFT2_R2>(kind, c, st)
              : launch<double, OA_FFT2_R1, OA_FFT2_R2>(kind, c, st);
}
}  // namespace oa2
