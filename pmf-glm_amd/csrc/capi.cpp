// C ABI of the MI355X GLM-HLLD step (declared in include/pmfmhd.h).
//
// Host-side mirror of lib/MHDKernels: allocate_state (state.jl:123-141), step! /
// evolve! (driver.jl:20-71), compute_dt! (timestep.jl:37-41), the diagnostics
// (diagnostics.jl, ct.jl:176-201) and the CT set-up (ct.jl:228-236,332-345). Errors
// reuse the reference's messages for the same misuse.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "state.hpp"

using namespace pmf;

namespace {
thread_local std::string g_err;
}  // namespace

void pmf::set_error(const std::string& msg) { g_err = msg; }

namespace {

double smax_device(pmf_state* s) {
  if (!s->smax_valid) {
    if (s->p.precision == PMF_F64)
      launch_wavespeed<double>(cfields_at<double>(s->U, s->nc), s->G, make_params<double>(s, 0, 0, 1), s->red,
                               s->dval, s->st);
    else
      launch_wavespeed<float>(cfields_at<float>(s->U, s->nc), s->G, make_params<float>(s, 0, 0, 1), s->red, s->dval,
                              s->st);
    hip_check(hipGetLastError(), "wavespeed launch");
    s->smax_valid = true;
  }
  fetch(s, 0, 1);
  return s->hval[0];
}

// One step from fields `in` into fields `out` (integrator_ref.jl:69-172 / the cube launch),
// then ct_update! on `out` for CT states (ct.jl:158-173). No swap.
template <class T>
void step_fields(pmf_state* s, Fields<const T> in, Fields<T> out, T* ctp, double dt, double ch, double decay,
                 int integrator) {
  const bool ct = s->p.solver == PMF_SOLVER_CT;
  bool edges = false;  // CT: the step kernel formed the edge EMFs
  if (integrator != PMF_INTEGRATOR_REF && integrator != PMF_INTEGRATOR_CUBE)
    throw Error("unknown integrator :" + std::to_string(integrator) + " (have :ref, :cube)");
  if (ct && integrator == PMF_INTEGRATOR_CUBE)
    throw Error("step!: CT is implemented on the :ref integrator only (mhd_solver=:ct states cannot use the :cube "
                "integrators)");
  if (!ct && !(ch > 0.0)) throw Error("step! needs the GLM cleaning speed `ch` (GLM states)");
  if (ct && !ctp) throw Error("ct_update! needs a CT state");
  if (integrator == PMF_INTEGRATOR_CUBE) {
    if (!fused_supported(s->G, s->p.precision, s->p.recon, s->p.riemann))
      throw Error("step_cube!: the fused kernel needs every axis at least one cell and reflecting axes two "
                  "(got " + std::to_string(s->p.nx) + "x" + std::to_string(s->p.ny) + "x" + std::to_string(s->p.nz) +
                  ")");
    mark(s);
    bool done =
        launch_step_fused<T>(in, out, s->G, make_params<T>(s, dt, ch, decay), s->red, s->dval, s->st, s->zc);
    mark(s);
    if (!done) throw Error("step_cube!: fused kernel unavailable for this configuration");
  } else if (ct && !s->ct_percell && fused_supported(s->G, s->p.precision, s->p.recon, s->p.riemann) &&
             all_periodic(s->G)) {
    // the same :ref algorithm with each face solved once (z-marching tiles); where the
    // variant allows, the edge EMFs of ct_update! come out of the same kernel
    mark(s);
    launch_step_fused_ct<T>(in, out, ctp, s->G, make_params<T>(s, dt, 0.0, 1.0), s->st, s->zc, nullptr,
                            !s->ct_fill_workspace, &edges);
    mark(s);
  } else {
    mark(s);
    launch_step_ref<T>(in, out, ctp, s->G, make_params<T>(s, dt, ct ? 0.0 : ch, ct ? 1.0 : decay), ct, s->st);
    mark(s);
  }
  hip_check(hipGetLastError(), "step launch");
  if (ct) {
    // ct_update! with the new state's CFL speed folded into dval[0] by its last pass
    const StepParams<T> Pw = make_params<T>(s, 0, 0, 1);
    launch_ct_update<T>(out, ctp, s->G, T(dt) / T(s->p.dx), false, s->st, &Pw, s->red, s->dval, edges);
    hip_check(hipGetLastError(), "ct_update launch");
  }
}

// The drag / radiation-coupled steps (step_ref_drag!, step_ref_radiation!,
// step_cube_drag!, step_cube_radiation!: integrator_ref.jl:222-258,
// integrator_cube.jl:478-540).
struct SourceCall {
  bool radiation = false;
  double impulse = 0.0;
  double v0[3] = {0.0, 0.0, 0.0};
  void* corr = nullptr;
  int flags = 0;  // PMF_SRC_* bits
};

// glm_physics.jl:63-65,87-103,118-122, evaluated in the state's precision.
template <class T> T drag_phi1(T q) {
  return std::fabs(q) < T(1e-4) ? T(1) - q / T(2) + q * q / T(6) : -std::expm1(-q) / q;
}
template <class T> void drag_work(T q, T& A, T& C, T& D) {
  if (q < T(0.05)) {
    T q2 = q * q, q3 = q2 * q, q4 = q3 * q;
    A = q * (T(1) - q + T(2.0 / 3.0) * q2 - T(1.0 / 3.0) * q3 + T(2.0 / 15.0) * q4);
    C = q * (T(1) - q + T(7.0 / 12.0) * q2 - T(1.0 / 4.0) * q3 + T(31.0 / 360.0) * q4);
    D = q * (T(1.0 / 3.0) - q / T(4) + T(7.0 / 60.0) * q2 - q3 / T(24) + T(31.0 / 2520.0) * q4);
    return;
  }
  T em1 = -std::expm1(-q), em2 = -std::expm1(-T(2) * q);
  A = T(0.5) * em2;
  C = em1 * em1 / q;
  D = (q - T(2) * em1 + T(0.5) * em2) / (q * q);
}

template <class T>
void step_fields_src(pmf_state* s, Fields<const T> in, Fields<T> out, T* ctp, double dt, double ch, double decay,
                     int integrator, const SourceCall& a) {
  const bool ct = s->p.solver == PMF_SOLVER_CT;
  const bool cube = integrator == PMF_INTEGRATOR_CUBE;
  const char* what = a.radiation ? (cube ? "step_cube_radiation!" : "step_ref_radiation!")
                                 : (cube ? "step_cube_drag!" : "step_ref_drag!");
  if (integrator != PMF_INTEGRATOR_REF && !cube)
    throw Error("unknown integrator :" + std::to_string(integrator) + " (have :ref, :cube)");
  if (ct && cube)
    throw Error(std::string(what) + ": CT is implemented on the :ref integrator only (mhd_solver=:ct states cannot "
                                    "use the :cube integrators)");
  if (cube && (a.v0[0] != 0.0 || a.v0[1] != 0.0 || a.v0[2] != 0.0))
    throw Error("step_cube_drag! is specialized for Compton drag in the CMB frame; use integrator=:ref for a moving "
                "target");
  const int checks = PMF_SRC_CHECK_ADMISSIBILITY | PMF_SRC_FALLBACK | PMF_SRC_GLOBAL_FALLBACK;
  if (!cube && (a.flags & checks))
    throw Error(std::string(what) + ": the admissibility check and the fallback masks are options of the :cube "
                                    "integrator");
  if ((a.flags & PMF_SRC_FALLBACK) && (a.flags & PMF_SRC_TRACK_DENSITY))
    throw Error(std::string(what) + ": the masked fallback reads correction slot 4, which density tracking writes");
  if ((a.radiation || (a.flags & PMF_SRC_FALLBACK)) && !a.corr)
    throw Error(std::string(what) + " needs the correction buffer (9 slots of ncells)");
  if (!ct && !(ch > 0.0)) throw Error("step! needs the GLM cleaning speed `ch` (GLM states)");
  if (ct && !ctp) throw Error("ct_update! needs a CT state");
  SrcParams<T> S{};
  const T q = T(a.impulse), qh = T(0.5) * q;
  S.face_decay = std::exp(-qh);
  S.face_phi1 = drag_phi1(qh);
  S.decay = std::exp(-q);
  S.phi1 = drag_phi1(q);
  drag_work(q, S.A, S.C, S.D);
  S.vx0 = T(a.v0[0]);
  S.vy0 = T(a.v0[1]);
  S.vz0 = T(a.v0[2]);
  S.corr = (T*)a.corr;
  S.stride = s->nc;
  S.flags = SRC_DRAG;
  if (a.radiation) {
    S.flags |= SRC_RAD;
    if (a.flags & PMF_SRC_TRACK_DENSITY) S.flags |= SRC_TRACK_RHO;
    if (a.flags & PMF_SRC_TRACK_MAGNETIC) S.flags |= SRC_TRACK_B;
    if (a.flags & PMF_SRC_TRACK_ENERGY) S.flags |= SRC_TRACK_E;
  }
  S.check = (a.flags & PMF_SRC_GLOBAL_FALLBACK)     ? CHECK_GLOBAL
            : (a.flags & PMF_SRC_FALLBACK)          ? CHECK_MASKED
            : (a.flags & PMF_SRC_CHECK_ADMISSIBILITY) ? CHECK_FLAG
                                                      : CHECK_NONE;
  // psi enters through slot 0 of the correction buffer (integrator_ref.jl:242-243)
  if (a.radiation) in.v[8] = (const T*)a.corr;
  StepParams<T> P = make_params<T>(s, dt, ct ? 0.0 : ch, ct ? 1.0 : decay);
  if (a.flags & PMF_SRC_ROBUST_FALLBACK) {  // the minmod-PLM / HLL trial (integrator_cube.jl:523-525)
    P.recon = RECON_PLM_MINMOD;
    P.rsolve = RS_HLL;
  } else if (a.flags & PMF_SRC_GLOBAL_FALLBACK) {
    P.rsolve = RS_LLF;
  }
  if (cube) {
    if (!fused_supported(s->G, s->p.precision, P.recon, P.rsolve))
      throw Error(std::string(what) + ": the fused kernel needs every axis at least one cell and reflecting axes two "
                                      "(got " + std::to_string(s->p.nx) + "x" + std::to_string(s->p.ny) + "x" +
                  std::to_string(s->p.nz) + ")");
    mark(s);
    bool done = launch_step_fused<T>(in, out, s->G, P, s->red, s->dval, s->st, s->zc, &S);
    mark(s);
    if (!done) throw Error(std::string(what) + ": fused kernel unavailable for this configuration");
  } else if (ct && !s->ct_percell && S.check != CHECK_MASKED && S.check != CHECK_GLOBAL &&
             fused_supported(s->G, s->p.precision, P.recon, P.rsolve) && all_periodic(s->G)) {
    // the CT :ref step with sources, each face solved once
    mark(s);
    launch_step_fused_ct<T>(in, out, ctp, s->G, P, s->st, s->zc, &S);
    mark(s);
  } else {
    mark(s);
    launch_step_ref<T>(in, out, ctp, s->G, P, ct, s->st, &S);
    mark(s);
  }
  hip_check(hipGetLastError(), "step launch");
  if (ct) {
    // compensated faces exactly when GLM would compensate cell B (integrator_ref.jl:255-256)
    const StepParams<T> Pw = make_params<T>(s, 0, 0, 1);
    launch_ct_update<T>(out, ctp, s->G, T(dt) / T(s->p.dx), a.radiation && (a.flags & PMF_SRC_TRACK_MAGNETIC), s->st,
                        &Pw, s->red, s->dval);
    hip_check(hipGetLastError(), "ct_update launch");
  }
}

void do_step_src(pmf_state* s, double dt, double ch, double decay, int integrator, const SourceCall& a) {
  check_fields(s);
  if (s->p.precision == PMF_F64)
    step_fields_src<double>(s, cfields_at<double>(s->U, s->nc), fields_at<double>(s->S, s->nc), (double*)s->ct, dt,
                            ch, decay, integrator, a);
  else
    step_fields_src<float>(s, cfields_at<float>(s->U, s->nc), fields_at<float>(s->S, s->nc), (float*)s->ct, dt, ch,
                           decay, integrator, a);
  std::swap(s->U, s->S);
  // the fused GLM epilogue or the CT cell-field pass reduced the new state's CFL speed
  s->smax_valid = (integrator == PMF_INTEGRATOR_CUBE && s->p.solver == PMF_SOLVER_GLM) || s->p.solver == PMF_SOLVER_CT;
}

// step_ref! / step_cube! on the state's own fields: into scratch, then swap.
void do_step(pmf_state* s, double dt, double ch, double decay, int integrator) {
  check_fields(s);
  if (s->p.precision == PMF_F64)
    step_fields<double>(s, cfields_at<double>(s->U, s->nc), fields_at<double>(s->S, s->nc), (double*)s->ct, dt, ch,
                        decay, integrator);
  else
    step_fields<float>(s, cfields_at<float>(s->U, s->nc), fields_at<float>(s->S, s->nc), (float*)s->ct, dt, ch,
                       decay, integrator);
  std::swap(s->U, s->S);
  // the fused GLM epilogue or the CT cell-field pass reduced the new state's CFL speed
  // into dval[0]
  s->smax_valid = (integrator == PMF_INTEGRATOR_CUBE && s->p.solver == PMF_SOLVER_GLM) || s->p.solver == PMF_SOLVER_CT;
}

}  // namespace

extern "C" {

int pmf_params_default(pmf_params* p, int32_t nx, int32_t ny, int32_t nz, double dx) {
  return guard([&] {
    if (!p) throw Error("null params");
    std::memset(p, 0, sizeof(*p));
    p->nx = nx;
    p->ny = ny;
    p->nz = nz;
    p->dx = dx;
    p->gamma = 5.0 / 3.0;
    p->smallr = 1e-6;
    p->pfl = 1e-7;
    p->llf_dmin = 1e-4;
    p->llf_pmin = 0.0;
    p->riemann = PMF_RIEMANN_HLLD;
    p->recon = PMF_RECON_PLM;
    p->solver = PMF_SOLVER_GLM;
    p->precision = PMF_F32;
  });
}

namespace {
// Parameter checks of allocate_state (state.jl:123-141) shared by every way of making a state.
void validate_params(const pmf_params& p) {
  if (p.nx < 1 || p.ny < 1 || p.nz < 1) throw Error("grid dimensions must be positive");
  if (!(p.dx > 0)) throw Error("dx must be positive");
  if (p.riemann < 0 || p.riemann > 2) throw Error("unknown MHD riemann code (have :hlld, :hll, :llf)");
  if (p.recon < 0 || p.recon > 2) throw Error("unknown recon code (have :plm, :plm_minmod, :ppm)");
  for (int d = 0; d < 3; ++d)
    if (p.bc[d] < 0 || p.bc[d] > 2) throw Error("unknown BC code (have :periodic, :outflow, :reflecting)");
  if (p.solver != PMF_SOLVER_GLM && p.solver != PMF_SOLVER_CT) throw Error("unknown mhd_solver (have :glm, :ct)");
  if (p.precision != PMF_F32 && p.precision != PMF_F64) throw Error("precision must be PMF_F32 or PMF_F64");
  if (p.solver == PMF_SOLVER_CT && (p.bc[0] || p.bc[1] || p.bc[2]))
    throw Error("mhd_solver=:ct requires all-periodic boundaries; constrained transport is periodic-only");
}

void create_state(const pmf_params* pp, pmf_state** out, bool with_fields) {
  if (!pp || !out) throw Error("null argument");
  const pmf_params& p = *pp;
  validate_params(p);
  auto* s = new pmf_state();
  try {
    s->p = p;
    s->nc = int64_t(p.nx) * p.ny * p.nz;
    s->G = Grid{p.nx, p.ny, p.nz, {p.bc[0], p.bc[1], p.bc[2]}, s->nc};
    s->esz = p.precision == PMF_F64 ? 8 : 4;
    hip_check(hipGetDevice(&s->device), "hipGetDevice");
    hip_check(hipStreamCreateWithFlags(&s->own, hipStreamNonBlocking), "hipStreamCreate");
    s->st = s->own;
    if (with_fields) {
      size_t bytes = 9 * s->nc * s->esz;
      hip_check(hipMalloc(&s->U, bytes), "hipMalloc(U)");
      hip_check(hipMalloc(&s->S, bytes), "hipMalloc(scratch)");
      hip_check(hipMemsetAsync(s->U, 0, bytes, s->st), "hipMemset");
      hip_check(hipMemsetAsync(s->S, 0, bytes, s->st), "hipMemset");
      if (p.solver == PMF_SOLVER_CT) {
        hip_check(hipMalloc(&s->ct, 21 * s->nc * s->esz), "hipMalloc(ct)");
        hip_check(hipMemsetAsync(s->ct, 0, 21 * s->nc * s->esz, s->st), "hipMemset");
      }
    }
    size_t nred = std::max<size_t>(size_t(kRedBlocks) * kRedSlots, size_t(fused_blocks(s->G, kMinChunk)));
    hip_check(hipMalloc(&s->red, sizeof(double) * nred), "hipMalloc(red)");
    hip_check(hipMalloc(&s->dval, sizeof(double) * 64), "hipMalloc(dval)");
    hip_check(hipHostMalloc((void**)&s->hval, sizeof(double) * 64, hipHostMallocDefault), "hipHostMalloc");
    sync(s);
  } catch (...) {
    pmf_destroy(s);
    throw;
  }
  *out = s;
}
}  // namespace

int pmf_create(const pmf_params* pp, pmf_state** out) {
  return guard([&] { create_state(pp, out, true); });
}

int pmf_create_workspace(const pmf_params* pp, pmf_state** out) {
  return guard([&] { create_state(pp, out, false); });
}

int pmf_destroy(pmf_state* s) {
  return guard([&] {
    if (!s) return;
    if (s->own) (void)hipStreamSynchronize(s->own);
    if (s->st && s->st != s->own) (void)hipStreamSynchronize(s->st);
    (void)hipFree(s->U);
    (void)hipFree(s->S);
    (void)hipFree(s->ct);
    (void)hipFree(s->red);
    (void)hipFree(s->dval);
    if (s->hval) (void)hipHostFree(s->hval);
    for (hipEvent_t e : s->ev) (void)hipEventDestroy(e);
    if (s->own) (void)hipStreamDestroy(s->own);
    delete s;
  });
}

int pmf_upload(pmf_state* s, int32_t var, const void* host, int64_t n) {
  return guard([&] {
    check_fields(s);
    if (var < 0 || var > 8) throw Error("variable index must be 0..8");
    if (n != s->nc) throw Error("host array has " + std::to_string(n) + " entries, expected " + std::to_string(s->nc));
    hip_check(hipMemcpyAsync((char*)s->U + var * s->nc * s->esz, host, n * s->esz, hipMemcpyHostToDevice, s->st),
              "upload");
    sync(s);
    s->smax_valid = false;
  });
}

int pmf_download(pmf_state* s, int32_t var, void* host, int64_t n) {
  return guard([&] {
    check_fields(s);
    if (var < 0 || var > 8) throw Error("variable index must be 0..8");
    if (n != s->nc) throw Error("host array has " + std::to_string(n) + " entries, expected " + std::to_string(s->nc));
    hip_check(hipMemcpyAsync(host, (char*)s->U + var * s->nc * s->esz, n * s->esz, hipMemcpyDeviceToHost, s->st),
              "download");
    sync(s);
  });
}

int pmf_device_ptr(pmf_state* s, int32_t var, void** ptr) {
  return guard([&] {
    check_fields(s);
    if (var < 0 || var > 8 || !ptr) throw Error("variable index must be 0..8");
    *ptr = (char*)s->U + var * s->nc * s->esz;
    s->smax_valid = false;  // the caller may write through the pointer
  });
}

int pmf_scratch_ptr(pmf_state* s, int32_t var, void** ptr) {
  return guard([&] {
    check_fields(s);
    if (var < 0 || var > 8 || !ptr) throw Error("variable index must be 0..8");
    *ptr = (char*)s->S + var * s->nc * s->esz;
  });
}

int pmf_ct_ptr(pmf_state* s, int32_t slot, void** ptr) {
  return guard([&] {
    check_state(s);
    if (!s->ct) throw Error("ct_update! needs a CT state");
    if (slot < 0 || slot > 20 || !ptr) throw Error("CT slot must be 0..20");
    *ptr = (char*)s->ct + slot * s->nc * s->esz;
    s->smax_valid = false;
  });
}

int pmf_ct_download(pmf_state* s, int32_t slot, void* host, int64_t n) {
  return guard([&] {
    check_state(s);
    if (!s->ct) throw Error("ct_update! needs a CT state");
    if (slot < 0 || slot > 20) throw Error("CT slot must be 0..20");
    if (n != s->nc) throw Error("host array has " + std::to_string(n) + " entries, expected " + std::to_string(s->nc));
    hip_check(hipMemcpyAsync(host, (char*)s->ct + slot * s->nc * s->esz, n * s->esz, hipMemcpyDeviceToHost, s->st),
              "download");
    sync(s);
  });
}

int pmf_set_stream(pmf_state* s, void* stream) {
  return guard([&] {
    check_state(s);
    sync(s);
    s->st = stream ? (hipStream_t)stream : s->own;
  });
}

int pmf_synchronize(pmf_state* s) {
  return guard([&] {
    check_state(s);
    sync(s);
  });
}

int pmf_max_wavespeed(pmf_state* s, double* smax) {
  return guard([&] {
    check_fields(s);
    *smax = smax_device(s);
  });
}

int pmf_compute_dt(pmf_state* s, double cfl, double* dt, double* smax) {
  return guard([&] {
    check_fields(s);
    double sm = smax_device(s);
    if (s->p.precision == PMF_F64) {
      *dt = cfl * s->p.dx / std::max(sm, std::numeric_limits<double>::epsilon());
    } else {
      float smf = float(sm);
      *dt = double(float(cfl) * float(s->p.dx) / std::max(smf, std::numeric_limits<float>::epsilon()));
    }
    if (smax) *smax = sm;
  });
}

int pmf_set_ct_kernel(pmf_state* s, int32_t per_cell) {
  return guard([&] {
    check_state(s);
    s->ct_percell = per_cell != 0;
  });
}

int pmf_set_ct_workspace(pmf_state* s, int32_t fill) {
  return guard([&] {
    check_state(s);
    s->ct_fill_workspace = fill != 0;
  });
}

int pmf_step_raw(pmf_state* s, double dt, double ch, double decay, int32_t integrator) {
  return guard([&] {
    check_state(s);
    do_step(s, dt, ch, decay, integrator);
  });
}

int pmf_step(pmf_state* s, double dt, double ch, double glm_cr, int32_t integrator) {
  return guard([&] {
    check_state(s);
    if (s->p.solver == PMF_SOLVER_CT) {
      do_step(s, dt, 0.0, 1.0, integrator);
      return;
    }
    if (!(ch > 0.0)) throw Error("step! needs the GLM cleaning speed `ch` (GLM states)");
    do_step(s, dt, ch, decay_of(s, dt, ch, glm_cr), integrator);
  });
}

int pmf_evolve(pmf_state* s, double tfinal, double cfl, double glm_ch_fac, double glm_cr, int32_t integrator,
               double* t_out, int64_t* n_out) {
  return guard([&] {
    check_state(s);
    if (!(glm_ch_fac > 0)) throw Error("glm_ch_fac must be positive");
    if (!(glm_cr >= 0)) throw Error("glm_cr must be non-negative");
    const bool ct = s->p.solver == PMF_SOLVER_CT;
    const bool f64 = s->p.precision == PMF_F64;
    double t = 0.0;
    int64_t n = 0;
    while (t < tfinal * (1 - 1e-9)) {
      double dt, smax;
      if (pmf_compute_dt(s, cfl, &dt, &smax)) throw Error(g_err);
      double ch, dtf;
      if (ct) {
        ch = 0.0;
        dtf = dt;
      } else if (f64) {
        ch = glm_ch_fac * smax;
        dtf = dt * smax / std::max(smax, ch);
      } else {
        float chf = float(glm_ch_fac) * float(smax);
        ch = chf;
        dtf = dt * double(float(smax)) / double(std::max(float(smax), chf));
      }
      if (t + dtf > tfinal) dtf = tfinal - t;
      if (ct) do_step(s, dtf, 0.0, 1.0, integrator);
      else do_step(s, dtf, ch, decay_of(s, dtf, ch, glm_cr), integrator);
      t += dtf;
      ++n;
    }
    if (t_out) *t_out = t;
    if (n_out) *n_out = n;
  });
}

extern "C++" {
namespace {
// Device scalar block (64 doubles): [0] CFL speed, [1..3] div B, [4..11] totals,
// [12] face div B, [16..47] the device-resident step scalars of pmf_advance, [48..49]
// its clock (elapsed time, last dt).
constexpr int kAdvParams = 16, kAdvClock = 48;
static_assert(sizeof(StepParams<double>) <= (kAdvClock - kAdvParams) * sizeof(double), "step scalars fit");

template <class T>
void advance_steps(pmf_state* s, int64_t nsteps, double cfl, double fac, double cr) {
  auto* Pd = reinterpret_cast<StepParams<T>*>(s->dval + kAdvParams);
  double* clock = s->dval + kAdvClock;
  const StepParams<T> base = make_params<T>(s, 0, 0, 1);
  if (!s->smax_valid) {
    launch_wavespeed<T>(cfields_at<T>(s->U, s->nc), s->G, base, s->red, s->dval, s->st);
    hip_check(hipGetLastError(), "wavespeed launch");
  }
  for (int64_t k = 0; k < nsteps; ++k) {
    launch_step_params<T>(s->dval, base, Pd, clock, cfl, fac, cr, s->p.dx, k == 0, s->st);
    mark(s);
    const bool done = launch_step_fused<T>(cfields_at<T>(s->U, s->nc), fields_at<T>(s->S, s->nc), s->G, base, s->red,
                                           s->dval, s->st, s->zc, nullptr, Pd);
    mark(s);
    if (!done) throw Error("pmf_advance: fused kernel unavailable for this configuration");
    hip_check(hipGetLastError(), "step launch");
    std::swap(s->U, s->S);
  }
  s->smax_valid = true;  // the last step folded its new state's speed into dval[0]
}
}  // namespace
}  // extern "C++"

int pmf_advance(pmf_state* s, int64_t nsteps, double cfl, double glm_ch_fac, double glm_cr, double* t_out,
                double* dt_out) {
  return guard([&] {
    check_fields(s);
    if (s->p.solver != PMF_SOLVER_GLM)
      throw Error("pmf_advance runs the fused GLM step; constrained-transport states step with pmf_step");
    if (nsteps < 0) throw Error("the number of steps must be non-negative");
    if (!(glm_ch_fac > 0)) throw Error("glm_ch_fac must be positive");
    if (!(glm_cr >= 0)) throw Error("glm_cr must be non-negative");
    if (!(cfl > 0)) throw Error("cfl must be positive");
    if (!fused_supported(s->G, s->p.precision, s->p.recon, s->p.riemann))
      throw Error("step_cube!: the fused kernel needs every axis at least one cell and reflecting axes two (got " +
                  std::to_string(s->p.nx) + "x" + std::to_string(s->p.ny) + "x" + std::to_string(s->p.nz) + ")");
    if (nsteps == 0) {
      if (t_out) *t_out = 0.0;
      if (dt_out) *dt_out = 0.0;
      return;
    }
    if (s->p.precision == PMF_F64) advance_steps<double>(s, nsteps, cfl, glm_ch_fac, glm_cr);
    else advance_steps<float>(s, nsteps, cfl, glm_ch_fac, glm_cr);
    fetch(s, kAdvClock, 2);
    if (t_out) *t_out = s->hval[kAdvClock];
    if (dt_out) *dt_out = s->hval[kAdvClock + 1];
  });
}

int pmf_set_chunk(pmf_state* s, int32_t zc) {
  return guard([&] {
    check_state(s);
    if (zc != 0 && zc < kMinChunk) throw Error("z chunks must be at least " + std::to_string(kMinChunk) +
                                               " planes (or 0 for the built-in choice)");
    s->zc = zc;
  });
}

int pmf_autotune(pmf_state* s, int32_t reps, int32_t* zc_out, double* ms_out) {
  return guard([&] {
    check_fields(s);
    if (reps < 1) throw Error("reps must be positive");
    if (!fused_supported(s->G, s->p.precision, s->p.recon, s->p.riemann))
      throw Error("autotune: the fused kernel does not run on this state");
    const bool ct = s->p.solver == PMF_SOLVER_CT;
    if (ct && !s->ct) throw Error("ct_update! needs a CT state");
    double dt, smax;
    if (pmf_compute_dt(s, 0.2, &dt, &smax)) throw Error(g_err);
    const double ch = smax > 0 ? smax : 1.0;
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    const int keep = s->zc;
    int best_zc = 0;
    double best = 1e300;
    try {
      for (int zc : {4, 8, 12, 16, 24, 32, 48, 64, 96, 128}) {
        if (zc > s->G.nz && zc != 4) break;
        s->zc = zc;
        // writes only the scratch fields (and, for CT, workspace slots 3-14 or the edge
        // EMFs 15-17, which every step recomputes): the state and its CT faces are untouched
        auto launch = [&]() {
          if (ct && s->p.precision == PMF_F64)
            launch_step_fused_ct<double>(cfields_at<double>(s->U, s->nc), fields_at<double>(s->S, s->nc),
                                         (double*)s->ct, s->G, make_params<double>(s, dt, 0.0, 1.0), s->st, s->zc,
                                         nullptr, !s->ct_fill_workspace);
          else if (ct)
            launch_step_fused_ct<float>(cfields_at<float>(s->U, s->nc), fields_at<float>(s->S, s->nc),
                                        (float*)s->ct, s->G, make_params<float>(s, dt, 0.0, 1.0), s->st, s->zc,
                                        nullptr, !s->ct_fill_workspace);
          else if (s->p.precision == PMF_F64)
            step_fields<double>(s, cfields_at<double>(s->U, s->nc), fields_at<double>(s->S, s->nc), nullptr, dt, ch,
                                1.0, PMF_INTEGRATOR_CUBE);
          else
            step_fields<float>(s, cfields_at<float>(s->U, s->nc), fields_at<float>(s->S, s->nc), nullptr, dt, ch, 1.0,
                               PMF_INTEGRATOR_CUBE);
          hip_check(hipGetLastError(), "autotune launch");
        };
        launch();  // warm-up
        hip_check(hipEventRecord(e0, s->st), "hipEventRecord");
        for (int r = 0; r < reps; ++r) launch();
        hip_check(hipEventRecord(e1, s->st), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        if (ms / reps < best) {
          best = ms / reps;
          best_zc = zc;
        }
      }
    } catch (...) {
      s->zc = keep;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    s->zc = best_zc;
    s->smax_valid = false;  // dval now holds the speed of the scratch fields
    if (zc_out) *zc_out = best_zc;
    if (ms_out) *ms_out = best;
  });
}

int pmf_max_divb(pmf_state* s, double* out) {
  return guard([&] {
    check_state(s);
    double l2, linf;
    if (pmf_divb_norm_fp64(s, &l2, &linf)) throw Error(g_err);
    *out = linf;
  });
}

int pmf_divb_norm_fp64(pmf_state* s, double* l2, double* linf) {
  return guard([&] {
    check_fields(s);
    if (s->p.precision == PMF_F64)
      launch_divb<double>(cfields_at<double>(s->U, s->nc), s->G, 1.0 / (2.0 * s->p.dx), s->red, s->dval + 1, s->st);
    else
      launch_divb<float>(cfields_at<float>(s->U, s->nc), s->G, 1.0f / (2.0f * float(s->p.dx)), s->red, s->dval + 1,
                         s->st);
    hip_check(hipGetLastError(), "divb launch");
    fetch(s, 1, 2);
    if (l2) *l2 = std::sqrt(s->hval[1] / double(s->nc));
    if (linf) *linf = s->hval[2];
  });
}

int pmf_conserved_totals_fp64(pmf_state* s, double out[8]) {
  return guard([&] {
    check_fields(s);
    if (s->p.precision == PMF_F64) launch_totals<double>(cfields_at<double>(s->U, s->nc), s->G, s->red, s->dval + 4, s->st);
    else launch_totals<float>(cfields_at<float>(s->U, s->nc), s->G, s->red, s->dval + 4, s->st);
    hip_check(hipGetLastError(), "totals launch");
    fetch(s, 4, 8);
    double dV = s->p.dx * s->p.dx * s->p.dx;
    for (int v = 0; v < 8; ++v) out[v] = s->hval[4 + v] * dV;
  });
}

int pmf_max_divb_face(pmf_state* s, double* out) {
  return guard([&] {
    check_state(s);
    if (!s->ct) throw Error("max_divb_face needs a CT state (mhd_solver=:ct)");
    if (s->p.precision == PMF_F64)
      launch_divb_face<double>((const double*)s->ct, s->G, 1.0 / s->p.dx, s->red, s->dval + 12, s->st);
    else
      launch_divb_face<float>((const float*)s->ct, s->G, 1.0f / float(s->p.dx), s->red, s->dval + 12, s->st);
    hip_check(hipGetLastError(), "divb_face launch");
    fetch(s, 12, 1);
    *out = s->hval[12];
  });
}

int pmf_ct_install_faces(pmf_state* s, const void* bx, const void* by, const void* bz) {
  return guard([&] {
    check_fields(s);
    if (!s->ct) throw Error("_ct_install_faces! needs a CT state (mhd_solver=:ct)");
    size_t b = s->nc * s->esz;
    const void* src[3] = {bx, by, bz};
    for (int d = 0; d < 3; ++d)
      hip_check(hipMemcpyAsync((char*)s->ct + d * b, src[d], b, hipMemcpyHostToDevice, s->st), "upload faces");
    if (s->p.precision == PMF_F64) launch_ct_install<double>(fields_at<double>(s->U, s->nc), (double*)s->ct, s->G, s->st);
    else launch_ct_install<float>(fields_at<float>(s->U, s->nc), (float*)s->ct, s->G, s->st);
    hip_check(hipGetLastError(), "ct_install launch");
    sync(s);
    s->smax_valid = false;
  });
}

int pmf_ct_faces_from_potential(pmf_state* s, const void* ax, const void* ay, const void* az) {
  return guard([&] {
    check_fields(s);
    if (!s->ct) throw Error("ct_faces_from_potential! needs a CT state (mhd_solver=:ct)");
    size_t b = s->nc * s->esz;
    // the potential is staged in scratch slots 0..2 (caller workspace between steps)
    const void* src[3] = {ax, ay, az};
    for (int d = 0; d < 3; ++d)
      hip_check(hipMemcpyAsync((char*)s->S + d * b, src[d], b, hipMemcpyHostToDevice, s->st), "upload potential");
    char* S = (char*)s->S;
    char* F = (char*)s->ct;
    if (s->p.precision == PMF_F64)
      launch_ct_curl<double>((double*)S, (double*)(S + b), (double*)(S + 2 * b), (double*)F, (double*)(F + b),
                             (double*)(F + 2 * b), s->G, 1.0 / s->p.dx, s->st);
    else
      launch_ct_curl<float>((float*)S, (float*)(S + b), (float*)(S + 2 * b), (float*)F, (float*)(F + b),
                            (float*)(F + 2 * b), s->G, 1.0f / float(s->p.dx), s->st);
    if (s->p.precision == PMF_F64) launch_ct_install<double>(fields_at<double>(s->U, s->nc), (double*)s->ct, s->G, s->st);
    else launch_ct_install<float>(fields_at<float>(s->U, s->nc), (float*)s->ct, s->G, s->st);
    hip_check(hipGetLastError(), "ct curl launch");
    sync(s);
    s->smax_valid = false;
  });
}

int pmf_riemann_flux_batch(pmf_state* s, const void* L, const void* R, int32_t dir, double ch, int64_t n, void* F) {
  return guard([&] {
    check_state(s);
    if (dir < 0 || dir > 2) throw Error("dir must be 0, 1 or 2");
    if (n < 0) throw Error("n must be non-negative");
    if (n == 0) return;
    const bool ct = s->p.solver == PMF_SOLVER_CT;
    if (!ct && !(ch > 0.0)) throw Error("riemann needs the GLM cleaning speed `ch` (GLM states)");
    size_t b = 9 * n * s->esz;
    void *dL = nullptr, *dR = nullptr, *dF = nullptr;
    try {
      hip_check(hipMalloc(&dL, b), "hipMalloc");
      hip_check(hipMalloc(&dR, b), "hipMalloc");
      hip_check(hipMalloc(&dF, b), "hipMalloc");
      hip_check(hipMemcpyAsync(dL, L, b, hipMemcpyHostToDevice, s->st), "copy");
      hip_check(hipMemcpyAsync(dR, R, b, hipMemcpyHostToDevice, s->st), "copy");
      if (s->p.precision == PMF_F64)
        launch_riemann_batch<double>((const double*)dL, (const double*)dR, (double*)dF, n, dir, ct,
                                     make_params<double>(s, 0, ct ? 0.0 : ch, 1), s->st);
      else
        launch_riemann_batch<float>((const float*)dL, (const float*)dR, (float*)dF, n, dir, ct,
                                    make_params<float>(s, 0, ct ? 0.0 : ch, 1), s->st);
      hip_check(hipGetLastError(), "riemann launch");
      hip_check(hipMemcpyAsync(F, dF, b, hipMemcpyDeviceToHost, s->st), "copy");
      sync(s);
    } catch (...) {
      (void)hipFree(dL);
      (void)hipFree(dR);
      (void)hipFree(dF);
      throw;
    }
    (void)hipFree(dL);
    (void)hipFree(dR);
    (void)hipFree(dF);
  });
}

int pmf_timing_enable(pmf_state* s, int32_t on) {
  return guard([&] {
    check_state(s);
    sync(s);
    s->timing = on != 0;
    s->ev_used = 0;
  });
}

int pmf_timing_read(pmf_state* s, double* total_ms, int64_t* launches) {
  return guard([&] {
    check_state(s);
    sync(s);
    double tot = 0.0;
    for (size_t i = 0; i + 1 < s->ev_used; i += 2) {
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, s->ev[i], s->ev[i + 1]), "hipEventElapsedTime");
      tot += ms;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = int64_t(s->ev_used / 2);
  });
}

int pmf_step_into(pmf_state* s, const void* const* U, void* const* O, void* ct, double dt, double ch, double decay,
                  int32_t integrator) {
  return guard([&] {
    check_state(s);
    if (!U || !O) throw Error("null field array");
    s->smax_valid = false;  // dval[0] now describes the caller's output arrays
    if (s->p.precision == PMF_F64)
      step_fields<double>(s, cfields_from<double>(U), fields_from<double>(O), (double*)ct, dt, ch, decay, integrator);
    else
      step_fields<float>(s, cfields_from<float>(U), fields_from<float>(O), (float*)ct, dt, ch, decay, integrator);
  });
}

int pmf_max_wavespeed_of(pmf_state* s, const void* const* U, double* smax) {
  return guard([&] {
    check_state(s);
    if (!U || !smax) throw Error("null argument");
    s->smax_valid = false;  // dval[0] now describes the caller's arrays
    if (s->p.precision == PMF_F64)
      launch_wavespeed<double>(cfields_from<double>(U), s->G, make_params<double>(s, 0, 0, 1), s->red, s->dval, s->st);
    else
      launch_wavespeed<float>(cfields_from<float>(U), s->G, make_params<float>(s, 0, 0, 1), s->red, s->dval, s->st);
    hip_check(hipGetLastError(), "wavespeed launch");
    fetch(s, 0, 1);
    *smax = s->hval[0];
  });
}

// ---- The allocation-free step --------------------------------------------------------
// pmf_step_into for callers that may not allocate device memory (code under HIP graph
// capture, serving loops, hosts that forbid it): no pmf_state, the caller's stream, and
// the kernels' workspace (per-workgroup partial maxima of the CFL fold, then the device
// scalars) in one file-scope device array. Every block writes its own partial slot before
// the final fold reads it, so nothing in the array has to start at zero. Calls that
// overlap on one device share the array and race.
namespace {
constexpr int64_t kStaticPartials = int64_t(1) << 18;  // doubles: per-workgroup partials
constexpr int64_t kStaticScalars = 64;                 // doubles: [0] = CFL speed of the output
constexpr int kStaticMaxDevices = 64;
}  // namespace

// 2 MiB + 512 B on each device that runs pmf_step_into_static, held until the process ends.
__device__ __align__(256) double scratch[kStaticPartials + kStaticScalars];

namespace {
std::mutex g_static_mu;
double* g_static_ptr[kStaticMaxDevices] = {};

// Device address of this device's copy of `scratch`, looked up once per device with that
// device current. Host code uses only this address, never the array's name.
double* static_workspace() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kStaticMaxDevices)
    throw Error("pmf_step_into_static: device ordinal " + std::to_string(dev) + " beyond the " +
                std::to_string(kStaticMaxDevices) + " the library keeps workspace addresses for");
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (!g_static_ptr[dev]) {
    void* ptr = nullptr;
    hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) throw Error("hipGetSymbolAddress(scratch) returned a null address");
    g_static_ptr[dev] = (double*)ptr;
  }
  return g_static_ptr[dev];
}

// Doubles of partials the step of pmf_step_into needs for this grid: the fused GLM kernel
// writes one per workgroup at the built-in chunk; the whole-grid reductions (CT's
// cell-field pass, the CFL reduction of the output that smax_dev asks for after a
// per-cell GLM step) use kRedSlots per block for at most kRedBlocks blocks.
int64_t static_partials_needed(const pmf_params& p, int integrator, bool want_smax) {
  const int64_t nc = int64_t(p.nx) * p.ny * p.nz;
  const Grid G{p.nx, p.ny, p.nz, {p.bc[0], p.bc[1], p.bc[2]}, nc};
  if (p.solver == PMF_SOLVER_GLM && integrator == PMF_INTEGRATOR_CUBE) return fused_blocks(G, 0);
  if (p.solver == PMF_SOLVER_CT || want_smax) return int64_t(kRedBlocks) * kRedSlots;
  return 0;
}
}  // namespace

int pmf_step_into_static(const pmf_params* p, const void* const* U, void* const* O, void* ct, double dt, double ch,
                         double decay, int32_t integrator, void* hip_stream, double* smax_dev) {
  return guard([&] {
    if (!p) throw Error("null argument");
    validate_params(*p);
    if (!U || !O) throw Error("null field array");
    const int64_t need = static_partials_needed(*p, integrator, smax_dev != nullptr);
    if (need > kStaticPartials)
      throw Error("pmf_step_into_static: a " + std::to_string(p->nx) + "x" + std::to_string(p->ny) + "x" +
                  std::to_string(p->nz) + " grid needs " + std::to_string(need) +
                  " partial slots, more than the static workspace's " + std::to_string(kStaticPartials) +
                  "; use pmf_step_into");
    double* w = static_workspace();
    pmf_state s;  // parameters, the caller's stream and the static workspace; owns nothing
    s.p = *p;
    s.nc = int64_t(p->nx) * p->ny * p->nz;
    s.G = Grid{p->nx, p->ny, p->nz, {p->bc[0], p->bc[1], p->bc[2]}, s.nc};
    s.esz = p->precision == PMF_F64 ? 8 : 4;
    s.st = (hipStream_t)hip_stream;
    s.red = w;
    s.dval = w + kStaticPartials;
    const bool folded = p->solver == PMF_SOLVER_CT || integrator == PMF_INTEGRATOR_CUBE;
    if (p->precision == PMF_F64) {
      step_fields<double>(&s, cfields_from<double>(U), fields_from<double>(O), (double*)ct, dt, ch, decay, integrator);
      if (smax_dev && !folded)
        launch_wavespeed<double>(cfields_from<double>((const void* const*)O), s.G, make_params<double>(&s, 0, 0, 1),
                                 s.red, s.dval, s.st);
    } else {
      step_fields<float>(&s, cfields_from<float>(U), fields_from<float>(O), (float*)ct, dt, ch, decay, integrator);
      if (smax_dev && !folded)
        launch_wavespeed<float>(cfields_from<float>((const void* const*)O), s.G, make_params<float>(&s, 0, 0, 1),
                                s.red, s.dval, s.st);
    }
    hip_check(hipGetLastError(), "static step launch");
    if (smax_dev)
      hip_check(hipMemcpyAsync(smax_dev, s.dval, sizeof(double), hipMemcpyDeviceToDevice, s.st),
                "hipMemcpyAsync(smax)");
  });
}

int pmf_step_drag(pmf_state* s, double dt, double ch, double decay, double drag_impulse,
                  const double* target_velocity, int32_t integrator) {
  return guard([&] {
    check_state(s);
    SourceCall a;
    a.impulse = drag_impulse;
    if (target_velocity)
      for (int d = 0; d < 3; ++d) a.v0[d] = target_velocity[d];
    do_step_src(s, dt, ch, decay, integrator, a);
  });
}

int pmf_step_radiation(pmf_state* s, double dt, double ch, double decay, double drag_impulse, void* correction,
                       int32_t flags, int32_t integrator) {
  return guard([&] {
    check_state(s);
    SourceCall a;
    a.radiation = true;
    a.impulse = drag_impulse;
    a.corr = correction;
    a.flags = flags;
    do_step_src(s, dt, ch, decay, integrator, a);
  });
}

int pmf_step_sources_into(pmf_state* s, const void* const* U, void* const* O, void* ct, void* correction, double dt,
                          double ch, double decay, double drag_impulse, const double* target_velocity,
                          int32_t flags, int32_t integrator) {
  return guard([&] {
    check_state(s);
    if (!U || !O) throw Error("null field array");
    s->smax_valid = false;  // the fused epilogue leaves dval[0] describing the caller's output arrays
    SourceCall a;
    a.radiation = correction != nullptr;
    a.impulse = drag_impulse;
    a.corr = correction;
    a.flags = flags;
    if (target_velocity)
      for (int d = 0; d < 3; ++d) a.v0[d] = target_velocity[d];
    if (s->p.precision == PMF_F64)
      step_fields_src<double>(s, cfields_from<double>(U), fields_from<double>(O), (double*)ct, dt, ch, decay,
                              integrator, a);
    else
      step_fields_src<float>(s, cfields_from<float>(U), fields_from<float>(O), (float*)ct, dt, ch, decay, integrator,
                             a);
  });
}

int pmf_species_to_fractions(pmf_state* s, void* species1, void* species2, const void* rho, double maximum_total) {
  return guard([&] {
    check_state(s);
    if (!species1) throw Error("null species field");
    if (!rho && !s->U) throw Error("pass the gas density (this state is a workspace without fields)");
    const void* r = rho ? rho : s->U;  // the density is the first field
    if (s->p.precision == PMF_F64)
      launch_species_to_fractions<double>((double*)species1, (double*)species2, (const double*)r, s->nc,
                                          maximum_total, s->st);
    else
      launch_species_to_fractions<float>((float*)species1, (float*)species2, (const float*)r, s->nc,
                                         float(maximum_total), s->st);
    hip_check(hipGetLastError(), "species launch");
    sync(s);
  });
}

namespace {
void backtrace_call(pmf_state* s, void* a1, void* a2, void* t1, void* t2, double dt, double maxt, int mode,
                    const char* what) {
  check_fields(s);
  if (!a1 || !t1 || (mode == 1 && (!a2 || !t2))) throw Error("null field or scratch");
  if (!(s->p.nx == s->p.ny && s->p.ny == s->p.nz))
    throw Error(std::string(what) + " remap currently requires a cubic grid");
  const int n = s->p.nx;
  const size_t bytes = size_t(s->nc) * s->esz;
  if (s->p.precision == PMF_F64)
    launch_backtrace<double>((double*)t1, (double*)t2, (const double*)a1, (const double*)a2,
                             cfields_at<double>(s->U, s->nc), n, dt / s->p.dx, maxt, mode, s->st);
  else
    launch_backtrace<float>((float*)t1, (float*)t2, (const float*)a1, (const float*)a2,
                            cfields_at<float>(s->U, s->nc), n, float(dt) / float(s->p.dx), float(maxt), mode, s->st);
  hip_check(hipGetLastError(), "backtrace launch");
  hip_check(hipMemcpyAsync(a1, t1, bytes, hipMemcpyDeviceToDevice, s->st), "copy");
  if (mode == 1) hip_check(hipMemcpyAsync(a2, t2, bytes, hipMemcpyDeviceToDevice, s->st), "copy");
  sync(s);
}
}  // namespace

int pmf_advect_species(pmf_state* s, void* species1, void* species2, void* scratch1, void* scratch2, double dt,
                       double maximum_total) {
  return guard([&] {
    check_state(s);
    backtrace_call(s, species1, species2, scratch1, scratch2, dt, maximum_total, species2 ? 1 : 0, "species");
  });
}

int pmf_advect_passive_scalar(pmf_state* s, void* scalar, void* scratch, double dt) {
  return guard([&] {
    check_state(s);
    backtrace_call(s, scalar, nullptr, scratch, nullptr, dt, 0.0, 2, "scalar");
  });
}

int pmf_buffer_alloc(int64_t bytes, void** out) {
  return guard([&] {
    if (!out || bytes <= 0) throw Error("bad buffer request");
    *out = nullptr;
    hip_check(hipMalloc(out, size_t(bytes)), "hipMalloc");
    hip_check(hipMemset(*out, 0, size_t(bytes)), "hipMemset");
    // the fill runs on the null stream, which the library's non-blocking streams do not wait for
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  });
}

int pmf_buffer_free(void* p) {
  return guard([&] {
    if (p) hip_check(hipFree(p), "hipFree");
  });
}

int pmf_buffer_copy(void* dst, const void* src, int64_t bytes, int32_t kind) {
  return guard([&] {
    if (!dst || !src || bytes < 0) throw Error("bad copy request");
    if (kind < 1 || kind > 3) throw Error("kind: 1 host to device, 2 device to host, 3 device to device");
    hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    hip_check(hipMemcpy(dst, src, size_t(bytes), k), "hipMemcpy");
    // hipMemcpy from pageable host memory (and device to device) may return before the data
    // has landed; work queued next on a non-blocking stream is not ordered after it
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  });
}

const char* pmf_last_error(void) { return g_err.c_str(); }

const char* pmf_version(void) { return "pmfmhd 0.1 (gfx950)"; }

}  // extern "C"
