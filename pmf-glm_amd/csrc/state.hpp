// Private host-side definitions shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pmfmhd.h"
#include "launch.hpp"

namespace pmf {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void set_error(const std::string& msg);

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class F> int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    set_error(e.what());
    return 1;
  } catch (...) {
    set_error("unknown error");
    return 1;
  }
}

}  // namespace pmf

struct pmf_state {
  pmf_params p;
  pmf::Grid G;
  int64_t nc;
  size_t esz;
  void* U = nullptr;     // 9*nc
  void* S = nullptr;     // 9*nc scratch
  void* ct = nullptr;    // 21*nc (CT only)
  double* red = nullptr; // reduction partials
  double* dval = nullptr;  // device scalars: [0]=smax [1..3]=divb [4..11]=totals [12]=face div
  double* hval = nullptr;  // pinned host mirror of dval
  hipStream_t own = nullptr;
  hipStream_t st = nullptr;
  int device = 0;
  bool smax_valid = false;  // dval[0] holds the CFL speed of the current U
  int zc = 0;               // planes per z chunk of the fused kernel (0 = built-in choice)
  bool ct_percell = false;  // CT steps on the per-cell kernel even where the fused one applies
  // CT steps on the fused kernel also write workspace slots 3-14 (face mass and field fluxes,
  // cell EMFs) and form the edge EMFs in a separate pass, as the reference's step does;
  // otherwise the fused kernel forms the edge EMFs itself and leaves slots 3-14 alone
  bool ct_fill_workspace = false;
  bool timing = false;      // bracket step kernels with events
  std::vector<hipEvent_t> ev;  // pairs (start, stop) in launch order
  size_t ev_used = 0;
};

namespace pmf {

inline void check_state(const pmf_state* s) {
  if (!s) throw Error("null pmf_state");
}

inline void check_fields(const pmf_state* s) {
  check_state(s);
  if (!s->U) throw Error("this pmf_state is a workspace without fields; use the *_into / *_of entry points");
}

template <class T> Fields<T> fields_at(void* base, int64_t nc) {
  Fields<T> f;
  for (int v = 0; v < 9; ++v) f.v[v] = (T*)base + v * nc;
  return f;
}
template <class T> Fields<const T> cfields_at(const void* base, int64_t nc) {
  Fields<const T> f;
  for (int v = 0; v < 9; ++v) f.v[v] = (const T*)base + v * nc;
  return f;
}
template <class T> Fields<T> fields_from(void* const* ptrs) {
  Fields<T> f;
  for (int v = 0; v < 9; ++v) {
    if (!ptrs[v]) throw Error("null field pointer");
    f.v[v] = (T*)ptrs[v];
  }
  return f;
}
template <class T> Fields<const T> cfields_from(const void* const* ptrs) {
  Fields<const T> f;
  for (int v = 0; v < 9; ++v) {
    if (!ptrs[v]) throw Error("null field pointer");
    f.v[v] = (const T*)ptrs[v];
  }
  return f;
}

template <class T> StepParams<T> make_params(const pmf_state* s, double dt, double ch, double decay) {
  const pmf_params& p = s->p;
  StepParams<T> P;
  P.gamma = T(p.gamma);
  P.gm1 = T(p.gamma) - T(1);
  P.igm1 = T(1) / P.gm1;
  P.smallr = T(p.smallr);
  P.pfl = T(p.pfl);
  P.llf_dmin = T(p.llf_dmin);
  P.llf_pmin = T(p.llf_pmin);
  P.ch = T(ch);
  P.ich = ch != 0.0 ? T(1) / T(ch) : T(0);
  P.ch2 = T(ch) * T(ch);
  P.dtdx = T(dt) / T(p.dx);
  P.hdtdx = T(0.5) * P.dtdx;
  P.decay = T(decay);
  P.recon = p.recon;
  P.rsolve = p.riemann;
  return P;
}


inline void sync(pmf_state* s) { hip_check(hipStreamSynchronize(s->st), "hipStreamSynchronize"); }

// Records one timing event on the state's stream (two per timed launch).
inline void mark(pmf_state* s) {
  if (!s->timing) return;
  if (s->ev_used == s->ev.size()) {
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    s->ev.push_back(e);
  }
  hip_check(hipEventRecord(s->ev[s->ev_used++], s->st), "hipEventRecord");
}

inline void fetch(pmf_state* s, int first, int count) {
  hip_check(hipMemcpyAsync(s->hval + first, s->dval + first, count * sizeof(double), hipMemcpyDeviceToHost, s->st),
            "hipMemcpyAsync");
  sync(s);
}


// GLM psi damping factor exp(-cr ch dt / dx), evaluated in the state's precision.
inline double decay_of(const pmf_state* s, double dt, double ch, double glm_cr) {
  if (s->p.precision == PMF_F64) return std::exp(-glm_cr * ch * dt / s->p.dx);
  // float32: exp evaluated in float64 and rounded once (the ensemble's device-side CFL
  // stage does the same, so both paths get the same factor)
  float v = -float(glm_cr) * float(ch) * float(dt) / float(s->p.dx);
  return double(float(std::exp(double(v))));
}

}  // namespace pmf
