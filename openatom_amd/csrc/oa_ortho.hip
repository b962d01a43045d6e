// oa_ortho.hip -- stateless T = (S/2)^{-1/2} on the device: the interface of
// CPORTHOG::CP_orthoTransform (cp_lowdin.C:24-46, class_cporthog.h:24) for
// callers that hold S themselves.
//
// It runs the Newton-Schulz iteration of get_iter_Tmat (cp_lowdin.C:180-265):
// X0 = S/2, T0 = I; C = 3I - T X, X' = X C/2, T' = C T/2 until the rms change
// of T is <= tol, on the same kernels as a context's oa_ortho_ns.  The work
// space (five n x n matrices, the per-workgroup error partials of every
// iteration and the device result record) is allocated per call, like the
// reference's scratch matrices new[]'d per call; S and T are host arrays
// (copy-in and copy-out).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "oa_kernels.h"
#include "oa_step.h"

namespace {

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of each part of the work space for n states and max_iter iterations.
struct NsWork {
  size_t mat;    // one n x n matrix (256-byte granules)
  size_t err;    // error partials: (max_iter + 2) x the largest workgroup count
  size_t res;    // StepDev
  size_t total;  // 5 mat + err + res
};

constexpr NsWork ns_work(int n, int max_iter) {
  const size_t nt = (size_t)(n + 31) / 32, nb = (size_t)(n + 15) / 16;
  const size_t nwg = std::max(nt * nt, (size_t)oa::kNsRowsMaxSplit * nb);
  NsWork w{};
  w.mat = align256((size_t)n * n * sizeof(double));
  w.err = align256((size_t)(max_iter + 2) * nwg * sizeof(double));
  w.res = align256(sizeof(oa::StepDev));
  w.total = 5 * w.mat + w.err + w.res;
  return w;
}

// Newton-Schulz arguments over a work space that starts at `base`.
oa::NsArgs ns_args(const double* S, double* T, unsigned char* base, int n, int max_iter,
                   double tol) {
  const NsWork w = ns_work(n, max_iter);
  oa::NsArgs a{};
  a.S = S;
  a.X0 = reinterpret_cast<double*>(base);
  a.X1 = reinterpret_cast<double*>(base + w.mat);
  a.T0 = reinterpret_cast<double*>(base + 2 * w.mat);
  a.T1 = reinterpret_cast<double*>(base + 3 * w.mat);
  a.C = reinterpret_cast<double*>(base + 4 * w.mat);
  a.errpart = reinterpret_cast<double*>(base + 5 * w.mat);
  a.res = reinterpret_cast<oa::StepDev*>(base + 5 * w.mat + w.err);
  a.Tout = T;
  a.n = n;
  a.max_iter = max_iter;
  a.tol = tol;
  return a;
}

// The fastest Newton-Schulz form that fits this device: one cooperative
// launch over 16-row blocks (split over 4, 2, 1 workgroups), over 32 x 32
// tiles, else two launches per iteration.
int ns_kind(int n, int device) {
  for (int kind = 4; kind >= 2; kind--)
    if (oa::ns_rows_possible(n, oa::ns_rows_split(kind), device)) return kind;
  return oa::ns_coop_possible(n, device) ? 1 : 0;
}

// Iterations used (> 0) or -max_iter, read back after the stream drains.
oa_status ns_finish(const oa::StepDev* res, hipStream_t st, int* iters) {
  int it = 0;
  if (hipMemcpyAsync(&it, &res->ns_iter, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return OA_EHIP;
  if (iters) *iters = it < 0 ? -it : it;
  return it > 0 ? OA_OK : OA_ENOCONV;
}

// get_iter_Tmat's own limits (cp_lowdin.C:212)
constexpr int kOrthoMaxIter = 10;
constexpr double kOrthoTol = 1e-15;

// ---- allocation-free variant: the work space, S and T live in a static
// device array sized for n <= OA_ORTHO_NOALLOC_MAX_N (14.7 MB per device,
// held until the process ends).  Nothing in it has to start at zero: the
// iteration writes X0, T0, every error partial and the result record before
// reading them.
constexpr size_t kNoAllocMat =
    align256((size_t)OA_ORTHO_NOALLOC_MAX_N * OA_ORTHO_NOALLOC_MAX_N * sizeof(double));
constexpr size_t kScratchBytes =
    ns_work(OA_ORTHO_NOALLOC_MAX_N, kOrthoMaxIter).total + 2 * kNoAllocMat;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Per device: the address of that device's copy of `scratch` (the array's
// name is a host-side stand-in, never passed to a kernel), a stream created
// once, and a lock that serialises this process's calls on the device.
struct ScratchDev {
  std::once_flag once;
  hipError_t rc = hipSuccess;
  unsigned char* ptr = nullptr;
  hipStream_t st = nullptr;
  std::mutex mu;
};
constexpr int kMaxDevices = 64;
ScratchDev g_scratch[kMaxDevices];

// set-up, once per device with that device current
void scratch_init(ScratchDev& d) {
  d.rc = hipGetSymbolAddress(reinterpret_cast<void**>(&d.ptr), HIP_SYMBOL(scratch));
  if (d.rc == hipSuccess && !d.ptr) d.rc = hipErrorInvalidSymbol;
  if (d.rc == hipSuccess) d.rc = hipStreamCreateWithFlags(&d.st, hipStreamNonBlocking);
}

}  // namespace

extern "C" {

oa_status oa_ortho_transform(double* T, const double* S, int n) {
  if (!T || !S || n < 1) return OA_EINVAL;
  const int max_iter = kOrthoMaxIter;
  const double tol = kOrthoTol;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return OA_EHIP;
  const NsWork w = ns_work(n, max_iter);
  const size_t nn = (size_t)n * n * sizeof(double);
  unsigned char* work = nullptr;
  double *dS = nullptr, *dT = nullptr;
  hipStream_t st = nullptr;
  oa_status s = OA_OK;
  if (hipMalloc(&work, w.total) != hipSuccess || hipMalloc(&dS, nn) != hipSuccess ||
      hipMalloc(&dT, nn) != hipSuccess) {
    (void)hipGetLastError();
    s = OA_ENOMEM;
  }
  if (s == OA_OK && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) s = OA_EHIP;
  if (s == OA_OK && hipMemcpyAsync(dS, S, nn, hipMemcpyHostToDevice, st) != hipSuccess) s = OA_EHIP;
  if (s == OA_OK) {
    oa::NsArgs a = ns_args(dS, dT, work, n, max_iter, tol);
    a.coop = ns_kind(n, dev);
    if (oa::ns_run(a, st) != hipSuccess) s = OA_EHIP;
    if (s == OA_OK) s = ns_finish(a.res, st, nullptr);
    // T is returned also when the iteration missed its tolerance
    if ((s == OA_OK || s == OA_ENOCONV) &&
        (hipMemcpyAsync(T, dT, nn, hipMemcpyDeviceToHost, st) != hipSuccess ||
         hipStreamSynchronize(st) != hipSuccess))
      s = OA_EHIP;
  }
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(work);
  (void)hipFree(dS);
  (void)hipFree(dT);
  return s;
}

oa_status oa_ortho_transform_noalloc(double* T, const double* S, int n) {
  if (!T || !S || n < 1 || n > OA_ORTHO_NOALLOC_MAX_N) return OA_EINVAL;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return OA_EHIP;
  if (dev < 0 || dev >= kMaxDevices) return OA_EINVAL;
  ScratchDev& d = g_scratch[dev];
  std::call_once(d.once, scratch_init, std::ref(d));
  if (d.rc != hipSuccess) return OA_EHIP;
  std::lock_guard<std::mutex> lk(d.mu);
  const NsWork w = ns_work(n, kOrthoMaxIter);
  const size_t nn = (size_t)n * n * sizeof(double);
  double* dS = reinterpret_cast<double*>(d.ptr + w.total);
  double* dT = reinterpret_cast<double*>(d.ptr + w.total + kNoAllocMat);
  if (hipMemcpyAsync(dS, S, nn, hipMemcpyHostToDevice, d.st) != hipSuccess) return OA_EHIP;
  oa::NsArgs a = ns_args(dS, dT, d.ptr, n, kOrthoMaxIter, kOrthoTol);
  a.coop = ns_kind(n, dev);
  if (oa::ns_run(a, d.st) != hipSuccess) return OA_EHIP;
  oa_status s = ns_finish(a.res, d.st, nullptr);
  if ((s == OA_OK || s == OA_ENOCONV) &&
      (hipMemcpyAsync(T, dT, nn, hipMemcpyDeviceToHost, d.st) != hipSuccess ||
       hipStreamSynchronize(d.st) != hipSuccess))
    s = OA_EHIP;
  return s;
}

}  // extern "C"
