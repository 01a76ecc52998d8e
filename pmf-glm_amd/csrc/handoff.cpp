// The production runner's raw state handoff (src/production/runner.jl:3814-3883): a
// directory of mhd_U_1.f32 ... mhd_U_9.f32, each the nx*ny*nz float32 values of one
// conserved variable in Julia's column-major order (x fastest), which is the library's
// own cell order. The runner writes these files with _write_f32_array_atomic
// (runner.jl:3814-3825: the raw bytes through a .tmp file and a rename) and reads them
// with _read_f32_array_into! and _import_initial_mhd_state! (runner.jl:3838-3883: exact
// size, then positive finite minimum density and energy). Errors reuse its messages.
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <limits>
#include <string>
#include <vector>

#include "state.hpp"

using namespace pmf;

namespace {

std::string var_path(const char* dir, int v) { return std::string(dir) + "/mhd_U_" + std::to_string(v + 1) + ".f32"; }

// Julia's print of a Float64 for the error texts (shortest round trip is not needed:
// the messages are matched by prefix).
std::string num(double x) {
  char b[64];
  std::snprintf(b, sizeof b, "%.9g", x);
  return b;
}

// _read_f32_array_into! (runner.jl:3838-3851): exactly n float32 values or the
// reference's size error (a missing file has size 0, as filesize reports it).
std::vector<float> read_f32(const std::string& path, int64_t n) {
  struct stat st {};
  const int64_t size = ::stat(path.c_str(), &st) == 0 ? int64_t(st.st_size) : 0;
  const int64_t want = n * int64_t(sizeof(float));
  if (size != want)
    throw Error("bad raw Float32 size for " + path + ": got " + std::to_string(size) + " bytes, expected " +
                std::to_string(want));
  std::vector<float> host(static_cast<size_t>(n));
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw Error("cannot open " + path);
  const size_t got = std::fread(host.data(), sizeof(float), host.size(), f);
  const bool trailing = std::fgetc(f) != EOF;
  std::fclose(f);
  if (got != host.size()) throw Error("short read of " + path);
  if (trailing) throw Error("extra trailing bytes in " + path);
  return host;
}

// _write_f32_array_atomic (runner.jl:3814-3825).
void write_f32(const std::string& path, const std::vector<float>& host) {
  const std::string tmp = path + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) throw Error("cannot create " + tmp);
  const size_t put = std::fwrite(host.data(), sizeof(float), host.size(), f);
  const bool ok = put == host.size() && std::fflush(f) == 0;
  std::fclose(f);
  if (!ok) {
    std::remove(tmp.c_str());
    throw Error("cannot write " + tmp);
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) throw Error("cannot rename " + tmp + " to " + path);
}

double minimum(const std::vector<float>& a) {
  double m = std::numeric_limits<double>::infinity();
  for (float x : a) {
    if (x != x) return double(x);  // Julia's minimum propagates NaN
    m = std::fmin(m, double(x));
  }
  return m;
}

}  // namespace

extern "C" int pmf_import_mhd_raw(pmf_state* s, const char* dir) {
  return guard([&] {
    check_fields(s);
    if (!dir) throw Error("null directory");
    std::vector<std::vector<float>> host(9);
    for (int v = 0; v < 9; ++v) host[v] = read_f32(var_path(dir, v), s->nc);
    // _import_initial_mhd_state! (runner.jl:3873-3878), checked before anything reaches
    // the device: a refused handoff leaves the state as it was
    const double rho_min = minimum(host[0]), energy_min = minimum(host[4]);
    if (!(std::isfinite(rho_min) && rho_min > 0))
      throw Error("imported MHD state has invalid minimum density " + num(rho_min));
    if (!(std::isfinite(energy_min) && energy_min > 0))
      throw Error("imported MHD state has invalid minimum energy " + num(energy_min));
    for (int v = 0; v < 9; ++v) {
      char* dst = (char*)s->U + v * s->nc * s->esz;
      if (s->p.precision == PMF_F64) {
        std::vector<double> d(host[v].begin(), host[v].end());  // copyto! converts (runner.jl:3847)
        hip_check(hipMemcpyAsync(dst, d.data(), d.size() * 8, hipMemcpyHostToDevice, s->st), "upload");
        sync(s);
      } else {
        hip_check(hipMemcpyAsync(dst, host[v].data(), host[v].size() * 4, hipMemcpyHostToDevice, s->st), "upload");
        sync(s);
      }
    }
    s->smax_valid = false;
  });
}

extern "C" int pmf_export_mhd_raw(pmf_state* s, const char* dir) {
  return guard([&] {
    check_fields(s);
    if (!dir) throw Error("null directory");
    std::vector<float> host(size_t(s->nc));
    for (int v = 0; v < 9; ++v) {
      const char* src = (const char*)s->U + v * s->nc * s->esz;
      if (s->p.precision == PMF_F64) {
        std::vector<double> d(size_t(s->nc));
        hip_check(hipMemcpyAsync(d.data(), src, d.size() * 8, hipMemcpyDeviceToHost, s->st), "download");
        sync(s);
        for (size_t i = 0; i < d.size(); ++i) host[i] = float(d[i]);  // Array{Float32} copyto! (runner.jl:3815-3816)
      } else {
        hip_check(hipMemcpyAsync(host.data(), src, host.size() * 4, hipMemcpyDeviceToHost, s->st), "download");
        sync(s);
      }
      write_f32(var_path(dir, v), host);
    }
  });
}
