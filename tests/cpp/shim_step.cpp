// One CP step through OaStepBackend, as a compiled C++ caller of the ABI
// would run it: stage by stage (gtoR, rtoG, lambda, integrate, ortho), then
// one whole step.  Inputs and outputs are raw little-endian files in the
// directory given as argv[1]:
//   in:  sys.txt "nx ny nz nstates npts ecut_ry dt", hmat.f64[9], kx/ky/kz.i32[npts],
//        psi.f64[nstates][npts][2], vks.f64[nz][ny][nx]
//   out: rho.f64, eke.txt, lam.f64, S.f64, T.f64, psi1.f64 (after the staged
//        step), step.txt "eelec fnorm ns_iters", psi2.f64 (after the whole step)
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "OaStepBackend.h"

void CkAbort(const char *msg) {
  std::fprintf(stderr, "CkAbort: %s\n", msg);
  std::exit(3);
}

template <typename T>
static std::vector<T> load(const std::string &path, size_t n) {
  std::vector<T> v(n);
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f || std::fread(v.data(), sizeof(T), n, f) != n) CkAbort(("cannot read " + path).c_str());
  std::fclose(f);
  return v;
}

template <typename T>
static void save(const std::string &path, const std::vector<T> &v) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f || std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size())
    CkAbort(("cannot write " + path).c_str());
  std::fclose(f);
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <dir>\n", argv[0]);
    return 2;
  }
  const std::string d = std::string(argv[1]) + "/";
  int grid[3], nst, npts;
  double ecut, dt;
  FILE *f = std::fopen((d + "sys.txt").c_str(), "r");
  if (!f || std::fscanf(f, "%d %d %d %d %d %lf %lf", &grid[0], &grid[1], &grid[2], &nst, &npts,
                        &ecut, &dt) != 7)
    CkAbort("cannot read sys.txt");
  std::fclose(f);
  const size_t N = (size_t)grid[0] * grid[1] * grid[2];
  const std::vector<double> hmat = load<double>(d + "hmat.f64", 9);
  const std::vector<int> kx = load<int>(d + "kx.i32", npts), ky = load<int>(d + "ky.i32", npts),
                         kz = load<int>(d + "kz.i32", npts);
  const std::vector<double> psi = load<double>(d + "psi.f64", (size_t)2 * nst * npts);
  const std::vector<double> vks = load<double>(d + "vks.f64", N);

  OaStepBackend be(hmat.data(), grid, ecut, nst, kx, ky, kz, dt, false);
  be.psiIn(psi.data());
  std::vector<double> rho(N), lam((size_t)nst * nst), S((size_t)nst * nst), T((size_t)nst * nst),
      out((size_t)2 * nst * npts);
  be.gtoR(rho.data());
  const double eke = be.rtoG(vks.data());
  be.lambda(lam.data());
  be.integrate();
  be.ortho(S.data(), T.data());
  be.psiOut(out.data());
  save(d + "rho.f64", rho);
  save(d + "lam.f64", lam);
  save(d + "S.f64", S);
  save(d + "T.f64", T);
  save(d + "psi1.f64", out);
  f = std::fopen((d + "eke.txt").c_str(), "w");
  std::fprintf(f, "%.17g\n", eke);
  std::fclose(f);
  oa_step_out so{};
  be.step(vks.data(), &so);
  be.psiOut(out.data());
  save(d + "psi2.f64", out);
  f = std::fopen((d + "step.txt").c_str(), "w");
  std::fprintf(f, "%.17g %.17g %d\n", so.eelec, so.fnorm, so.ns_iters);
  std::fclose(f);
  std::printf("shim: eke %.12f E %.12f |F| %.6e ns %d\n", eke, so.eelec, so.fnorm, so.ns_iters);
  return 0;
}
