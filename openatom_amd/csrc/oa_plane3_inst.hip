// One square plane size of the fused two-factor plane passes per translation
// unit: built with -DOA_FFT2_N=<N> -DOA_FFT2_R1=<R1> -DOA_FFT2_R2=<R2>
// (R1 <= R2).  The y transform runs as R1 x R2; the x transform as R2 x R1
// so that its pass 2 leaves R1 outputs per thread.
#include "oa_plane3.h"

#if !defined(OA_FFT2_N) || !defined(OA_FFT2_R1) || !defined(OA_FFT2_R2)
#error "OA_FFT2_N, OA_FFT2_R1 and OA_FFT2_R2 must be defined"
#endif
static_assert(OA_FFT2_R1 * OA_FFT2_R2 == OA_FFT2_N, "N must be R1 * R2");

#define OA_CAT2(a, b) a##b
#define OA_CAT(a, b) OA_CAT2(a, b)

namespace oa3 {
constexpr int kX1 = OA_FFT2_R2, kX2 = OA_FFT2_R1, kY1 = OA_FFT2_R1, kY2 = OA_FFT2_R2;

hipError_t OA_CAT(plane3_n, OA_FFT2_N)(int kind, bool fp32, const oa::FftCtx& c,
                                      hipStream_t st) {
  return fp32 ? launch_plane3<float, kX1, kX2, kY1, kY2>(kind, c, st)
              : launch_plane3<double, kX1, kX2, kY1, kY2>(kind, c, st);
}

size_t OA_CAT(plane3_smem_n, OA_FFT2_N)(int npx, bool fp32) {
  return fp32 ? plane3_smem<float, kX1, kX2, kY1, kY2, kPlaneThreads>(npx)
              : plane3_smem<double, kX1, kX2, kY1, kY2, kPlaneThreads>(npx);
}

size_t OA_CAT(plane3_kper_n, OA_FFT2_N)(int npx, bool fp32) {
  return fp32 ? plane3_kper<float, kX1, kX2, kY1, kY2>(npx)
              : plane3_kper<double, kX1, kX2, kY1, kY2>(npx);
}

hipError_t OA_CAT(plane3_info_n, OA_FFT2_N)(int npx, bool fp32, bool prune, int* rowparts,
                                           int* occ_g2r, int* occ_r2g, int* rho_ok) {
  return fp32 ? plane3_info<float, kX1, kX2, kY1, kY2>(npx, prune, rowparts, occ_g2r, occ_r2g,
                                                       rho_ok)
              : plane3_info<double, kX1, kX2, kY1, kY2>(npx, prune, rowparts, occ_g2r, occ_r2g,
                                                        rho_ok);
}
}  // namespace oa3
