// Host dispatch from an axis length to its compiled two-factor kernels.
#include "oa_fft2_sizes.h"
#include "oa_kernels.h"

namespace oa2 {
#define OA_DECL(n, r1, r2) \
  hipError_t launch_n##n(int kind, bool fp32, const oa::FftCtx& c, hipStream_t st);
OA_FFT2_SIZES(OA_DECL)
#undef OA_DECL
}  // namespace oa2

namespace oa3 {
#define OA_DECL(n, r1, r2)                                                                  \
  hipError_t plane3_n##n(int kind, bool fp32, const oa::FftCtx& c, hipStream_t st);        \
  size_t plane3_smem_n##n(int npx, bool fp32);                                              \
  size_t plane3_kper_n##n(int npx, bool fp32);                                              \
  hipError_t plane3_info_n##n(int npx, bool fp32, bool prune, int* rowparts, int* occ_g2r,    \
                              int* occ_r2g, int* rho_ok);
OA_FFT2_SIZES(OA_DECL)
#undef OA_DECL
}  // namespace oa3

namespace oa {

hipError_t fft2_launch(int kind, int n, bool fp32, const FftCtx& c, hipStream_t st) {
  switch (n) {
#define OA_CASE(nn, r1, r2) \
  case nn:                  \
    return oa2::launch_n##nn(kind, fp32, c, st);
    OA_FFT2_SIZES(OA_CASE)
#undef OA_CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t plane3_launch(int kind, int n, bool fp32, const FftCtx& c, hipStream_t st) {
  switch (n) {
#define OA_CASE(nn, r1, r2) \
  case nn:                  \
    return oa3::plane3_n##nn(kind, fp32, c, st);
    OA_FFT2_SIZES(OA_CASE)
#undef OA_CASE
    default:
      return hipErrorInvalidValue;
  }
}

size_t plane3_smem(int n, int npx, bool fp32) {
  switch (n) {
#define OA_CASE(nn, r1, r2) \
  case nn:                  \
    return oa3::plane3_smem_n##nn(npx, fp32);
    OA_FFT2_SIZES(OA_CASE)
#undef OA_CASE
    default:
      return 0;
  }
}

size_t plane3_kper(int n, int npx, bool fp32) {
  switch (n) {
#define OA_CASE(nn, r1, r2) \
  case nn:                  \
    return oa3::plane3_kper_n##nn(npx, fp32);
    OA_FFT2_SIZES(OA_CASE)
#undef OA_CASE
    default:
      return 0;
  }
}

hipError_t plane3_info(int n, int npx, bool fp32, bool prune, int* rowparts, int* occ_g2r,
                       int* occ_r2g, int* rho_ok) {
  switch (n) {
#define OA_CASE(nn, r1, r2) \
  case nn:                  \
    return oa3::plane3_info_n##nn(npx, fp32, prune, rowparts, occ_g2r, occ_r2g, rho_ok);
    OA_FFT2_SIZES(OA_CASE)
#undef OA_CASE
    default:
      return hipErrorInvalidValue;
  }
}

int fft2_ntr(int n) {
  switch (n) {
#define OA_CASE(nn, r1, r2) \
  case nn:                  \
    return 256 / ((r1) > (r2) ? (r1) : (r2));
    OA_FFT2_SIZES(OA_CASE)
#undef OA_CASE
    default:
      return 0;
  }
}

}  // namespace oa
