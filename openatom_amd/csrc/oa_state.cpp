// oa_state.cpp -- the reference's STATES files (include/oa_state.h): host
// code of the step library, no device work.
#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "oa_ctx_info.h"
#include "oa_state.h"

namespace {

uint64_t key3(int a, int b, int c) {
  const uint64_t off = 1u << 20;
  return (((uint64_t)(a + off)) << 42) | (((uint64_t)(b + off)) << 21) | (uint64_t)(c + off);
}

bool redundant(int kx, int ky, int kz) { return kx == 0 && (ky < 0 || (ky == 0 && kz < 0)); }

struct Rec {
  double re, im;
  int kx, ky, kz;
};

// Byte source over a plain or a gzip file (the ".gz" name is the caller's).
class Src {
 public:
  Src(const std::string& path, bool gz, bool binary) : gz_(gz) {
    if (gz)
      zf_ = gzopen(path.c_str(), "rb");
    else
      f_ = std::fopen(path.c_str(), binary ? "rb" : "r");
  }
  ~Src() {
    if (zf_) gzclose(zf_);
    if (f_) std::fclose(f_);
  }
  bool ok() const { return gz_ ? zf_ != nullptr : f_ != nullptr; }
  bool read(void* p, size_t n) {
    if (gz_) return gzread(zf_, p, (unsigned)n) == (int)n;
    return std::fread(p, 1, n, f_) == n;
  }
  // next line (text formats); false at end of file
  bool line(std::string& s) {
    char buf[1024];
    if (gz_) {
      if (!gzgets(zf_, buf, sizeof(buf))) return false;
    } else {
      if (!std::fgets(buf, sizeof(buf), f_)) return false;
    }
    s = buf;
    return true;
  }

 private:
  bool gz_;
  gzFile zf_ = nullptr;
  FILE* f_ = nullptr;
};

std::string file_name(const char* path, int format) {
  std::string p = path;
  if (format == OA_STATE_TEXT_GZ || format == OA_STATE_BINARY_GZ) p += ".gz";
  return p;
}

// readState (util.C:1107-1290): header and records in file order
bool read_records(const char* path, int format, int hdr[4], std::vector<Rec>& recs) {
  const bool gz = format == OA_STATE_TEXT_GZ || format == OA_STATE_BINARY_GZ;
  const bool binary = format == OA_STATE_BINARY || format == OA_STATE_BINARY_GZ;
  Src src(file_name(path, format), gz, binary);
  if (!src.ok()) return false;
  recs.clear();
  if (binary) {
    if (!src.read(hdr, 4 * sizeof(int)) || hdr[0] < 0) return false;
    recs.resize((size_t)hdr[0]);
    for (Rec& r : recs) {
      int k[3];
      if (!src.read(&r.re, sizeof(double)) || !src.read(&r.im, sizeof(double)) ||
          !src.read(k, sizeof(k)))
        return false;
      r.kx = k[0];
      r.ky = k[1];
      r.kz = k[2];
    }
    return true;
  }
  std::string s;
  if (!src.line(s) ||
      std::sscanf(s.c_str(), "%d%d%d%d", &hdr[0], &hdr[1], &hdr[2], &hdr[3]) != 4 || hdr[0] < 0)
    return false;
  recs.reserve((size_t)hdr[0]);
  for (int i = 0; i < hdr[0]; i++) {
    Rec r{};
    if (!src.line(s) ||
        std::sscanf(s.c_str(), "%lg%lg%d%d%d", &r.re, &r.im, &r.kx, &r.ky, &r.kz) != 5)
      return false;
    recs.push_back(r);
  }
  return true;
}

// Sink over a plain or a gzip file.
class Dst {
 public:
  Dst(const std::string& path, bool gz, bool binary) : gz_(gz) {
    if (gz)
      zf_ = gzopen(path.c_str(), "wb");
    else
      f_ = std::fopen(path.c_str(), binary ? "wb" : "w");
  }
  bool close() {
    bool ok = good_;
    if (zf_) ok = (gzclose(zf_) == Z_OK) && ok;
    if (f_) ok = (std::fclose(f_) == 0) && ok;
    zf_ = nullptr;
    f_ = nullptr;
    return ok;
  }
  ~Dst() { close(); }
  bool ok() const { return gz_ ? zf_ != nullptr : f_ != nullptr; }
  void write(const void* p, size_t n) {
    if (gz_)
      good_ = good_ && gzwrite(zf_, p, (unsigned)n) == (int)n;
    else
      good_ = good_ && std::fwrite(p, 1, n, f_) == n;
  }
  void text(const char* s) { write(s, std::strlen(s)); }

 private:
  bool gz_;
  bool good_ = true;
  gzFile zf_ = nullptr;
  FILE* f_ = nullptr;
};

}  // namespace

extern "C" {

oa_status oa_state_read(const char* path, int format, int npts, const int* kx, const int* ky,
                        const int* kz, double* psi, int grid[3]) {
  if (!path || format < 0 || format > 3 || npts < 1 || !kx || !ky || !kz || !psi)
    return OA_EINVAL;
  int hdr[4];
  std::vector<Rec> recs;
  if (!read_records(path, format, hdr, recs)) return OA_EINVAL;
  std::unordered_map<uint64_t, int> at;
  at.reserve((size_t)npts * 2);
  for (int i = 0; i < npts; i++) at[key3(kx[i], ky[i], kz[i])] = i;
  std::vector<char> seen((size_t)npts, 0);
  for (const Rec& r : recs) {
    const auto it = at.find(key3(r.kx, r.ky, r.kz));
    if (it == at.end()) return OA_EINVAL;  // a point outside the stored set
    psi[2 * (size_t)it->second] = r.re;
    psi[2 * (size_t)it->second + 1] = r.im;
    seen[(size_t)it->second] = 1;
  }
  // redundant kx = 0 half from the unique partners (flip_data_set)
  for (int i = 0; i < npts; i++) {
    if (!redundant(kx[i], ky[i], kz[i])) {
      if (!seen[(size_t)i]) return OA_EINVAL;  // a unique point missing from the file
      continue;
    }
    const auto it = at.find(key3(0, -ky[i], -kz[i]));
    if (it == at.end() || !seen[(size_t)it->second]) return OA_EINVAL;
    psi[2 * (size_t)i] = psi[2 * (size_t)it->second];
    psi[2 * (size_t)i + 1] = -psi[2 * (size_t)it->second + 1];
  }
  if (grid) {
    grid[0] = hdr[1];
    grid[1] = hdr[2];
    grid[2] = hdr[3];
  }
  return OA_OK;
}

oa_status oa_state_write(const char* path, int format, int npts, const int* kx, const int* ky,
                         const int* kz, const double* psi, const int grid[3]) {
  if (!path || format < 0 || format > 3 || npts < 1 || !kx || !ky || !kz || !psi || !grid)
    return OA_EINVAL;
  // unique points in (kx, ky, kz) order, G = 0 last (sort_psi_output and
  // the doublepack branch of writeStateFile)
  std::vector<int> idx;
  idx.reserve((size_t)npts);
  int g0 = -1;
  for (int i = 0; i < npts; i++) {
    if (redundant(kx[i], ky[i], kz[i])) continue;
    if (kx[i] == 0 && ky[i] == 0 && kz[i] == 0)
      g0 = i;
    else
      idx.push_back(i);
  }
  if (g0 < 0) return OA_EINVAL;
  std::sort(idx.begin(), idx.end(), [&](int a, int b) {
    if (kx[a] != kx[b]) return kx[a] < kx[b];
    if (ky[a] != ky[b]) return ky[a] < ky[b];
    return kz[a] < kz[b];
  });
  idx.push_back(g0);
  const bool gz = format == OA_STATE_TEXT_GZ || format == OA_STATE_BINARY_GZ;
  const bool binary = format == OA_STATE_BINARY || format == OA_STATE_BINARY_GZ;
  Dst dst(file_name(path, format), gz, binary);
  if (!dst.ok()) return OA_EINVAL;
  const int n = (int)idx.size();
  if (binary) {
    const int hdr[4] = {n, grid[0], grid[1], grid[2]};
    dst.write(hdr, sizeof(hdr));
    for (int i : idx) {
      dst.write(&psi[2 * (size_t)i], sizeof(double));
      dst.write(&psi[2 * (size_t)i + 1], sizeof(double));
      const int k[3] = {kx[i], ky[i], kz[i]};
      dst.write(k, sizeof(k));
    }
  } else {
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%d %d %d %d\n", n, grid[0], grid[1], grid[2]);
    dst.text(buf);
    for (int i : idx) {
      std::snprintf(buf, sizeof(buf), "%g %g %d %d %d \n", psi[2 * (size_t)i],
                    psi[2 * (size_t)i + 1], kx[i], ky[i], kz[i]);
      dst.text(buf);
    }
  }
  return dst.close() ? OA_OK : OA_EINVAL;
}

static std::string state_dir(const char* dir, int spin, int kpt, int bead, int temper) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "/Spin.%d_Kpt.%d_Bead.%d_Temper.%d", spin, kpt, bead, temper);
  return std::string(dir) + buf;
}

oa_status oa_states_read(oa_ctx* ctx, const char* dir, int format, int spin, int kpt, int bead,
                         int temper) {
  if (!ctx || !dir) return OA_EINVAL;
  const int nst = oa_ctx_nstates(ctx);
  int info[4];
  oa_status s = oa_layout_info(ctx, info);
  if (s != OA_OK) return s;
  const int npts = info[0];
  std::vector<int> kx(npts), ky(npts), kz(npts);
  if ((s = oa_get_kvecs(ctx, kx.data(), ky.data(), kz.data())) != OA_OK) return s;
  std::vector<double> psi((size_t)nst * npts * 2);
  const std::string d = state_dir(dir, spin, kpt, bead, temper);
  for (int i = 0; i < nst; i++) {
    const std::string p = d + "/state" + std::to_string(i + 1) + ".out";
    s = oa_state_read(p.c_str(), format, npts, kx.data(), ky.data(), kz.data(),
                      psi.data() + (size_t)i * npts * 2, nullptr);
    if (s != OA_OK) return s;
  }
  return oa_psi_set(ctx, psi.data());
}

oa_status oa_states_write(oa_ctx* ctx, const char* dir, int format, int spin, int kpt, int bead,
                          int temper, int iteration) {
  if (!ctx || !dir) return OA_EINVAL;
  const int nst = oa_ctx_nstates(ctx);
  int grid[3];
  oa_ctx_grid(ctx, grid);
  int info[4];
  oa_status s = oa_layout_info(ctx, info);
  if (s != OA_OK) return s;
  const int npts = info[0];
  std::vector<int> kx(npts), ky(npts), kz(npts);
  if ((s = oa_get_kvecs(ctx, kx.data(), ky.data(), kz.data())) != OA_OK) return s;
  std::vector<double> psi((size_t)nst * npts * 2);
  if ((s = oa_psi_get(ctx, psi.data())) != OA_OK) return s;
  const std::string d = state_dir(dir, spin, kpt, bead, temper);
  FILE* ts = std::fopen((d + "/TimeStamp").c_str(), "w");
  if (!ts) return OA_EINVAL;
  std::fprintf(ts, "time step = %d\n", iteration);
  if (std::fclose(ts) != 0) return OA_EINVAL;
  for (int i = 0; i < nst; i++) {
    const std::string p = d + "/state" + std::to_string(i + 1) + ".out";
    s = oa_state_write(p.c_str(), format, npts, kx.data(), ky.data(), kz.data(),
                       psi.data() + (size_t)i * npts * 2, grid);
    if (s != OA_OK) return s;
  }
  return OA_OK;
}

}  // extern "C"
