// Source terms of the drag / radiation-coupled steps: the half-step drag propagation of
// the face states and the radiation velocity correction (integrator_ref.jl:28-48), and
// the per-cell update with float residual tracking, the exact drag propagator and the
// admissibility mark (integrator_ref.jl:150-160, integrator_cube.jl:100-116,423-441,
// glm_physics.jl:143-170,350-390).
#pragma once
#include "mhd_device.hpp"

namespace pmf {

enum { SRC_DRAG = 1, SRC_RAD = 2, SRC_TRACK_RHO = 4, SRC_TRACK_B = 8, SRC_TRACK_E = 16 };
enum { CHECK_NONE = 0, CHECK_FLAG = 1, CHECK_MASKED = 2, CHECK_GLOBAL = 3 };

template <class T> struct SrcParams {
  T face_decay, face_phi1;  // exp(-q/2), phi1(q/2): the half-step face propagator
  T decay, phi1, A, C, D;   // the full-step propagator and the drag-work coefficients
  T vx0, vy0, vz0;          // drag target velocity
  T* corr;                  // radiation: 9 slots of `stride` cells (psi, dv_x, dv_y, dv_z,
                            // rho perturbation or fallback mask, B_x, B_y, B_z, E residuals)
  int64_t stride;
  int flags;  // SRC_* bits
  int check;  // CHECK_*: admissibility flag, masked fallback (slot 4 > 0), global fallback
};

// Face state with sources: q holds mh + d on entry, m0 + d is the state at the start of
// the step; idx is the predicted cell (its radiation correction).
template <class T>
__device__ __forceinline__ void src_face(T* q, const T* m0, const T* d, const SrcParams<T>& S, int64_t idx) {
  if (S.flags & SRC_DRAG) {
    const T v0[3] = {S.vx0, S.vy0, S.vz0};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T f0 = m0[1 + c] + d[1 + c];
      q[1 + c] = v0[c] + S.face_decay * (f0 - v0[c]) + S.face_phi1 * (q[1 + c] - f0);
    }
  }
  if (S.flags & SRC_RAD) {
#pragma unroll
    for (int c = 0; c < 3; ++c) q[1 + c] = q[1 + c] + S.corr[(1 + c) * S.stride + idx];
  }
}

// The same with the cell's radiation correction already loaded (cv = dv_x, dv_y, dv_z).
template <class T>
__device__ __forceinline__ void src_face_pre(T* q, const T* m0, const T* d, const SrcParams<T>& S, const T* cv) {
  if (S.flags & SRC_DRAG) {
    const T v0[3] = {S.vx0, S.vy0, S.vz0};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T f0 = m0[1 + c] + d[1 + c];
      q[1 + c] = v0[c] + S.face_decay * (f0 - v0[c]) + S.face_phi1 * (q[1 + c] - f0);
    }
  }
  if (S.flags & SRC_RAD) {
#pragma unroll
    for (int c = 0; c < 3; ++c) q[1 + c] = q[1 + c] + cv[c];
  }
}

// x + dx, carrying the rounding error in *res when tracked (Fast2Sum; glm_physics.jl:350-358).
template <class T> __device__ __forceinline__ T comp_update(T* res, T x, T dx, bool tracked) {
  if (!tracked) return x + dx;
  const T increment = dx + *res;
  const T updated = x + increment;
  *res = increment - (updated - x);
  return updated;
}

// integrator_cube.jl:100-116
template <class T> __device__ __forceinline__ bool src_inadmissible(const T* rh, const StepParams<T>& P) {
  const T rho = rh[0];
  const T gate = dmax(P.smallr, P.llf_dmin);
  const T ir = T(1) / dmax(rho, P.smallr);
  const T ek = T(0.5) * (rh[1] * rh[1] + rh[2] * rh[2] + rh[3] * rh[3]) * ir;
  const T eb = T(0.5) * (rh[5] * rh[5] + rh[6] * rh[6] + rh[7] * rh[7]);
  const T eint = rh[4] - ek - eb;
  const T emin = P.pfl / P.gm1;
  bool finite = __builtin_isfinite(rho) && __builtin_isfinite(eint);
#pragma unroll
  for (int v = 1; v < 9; ++v) finite = finite && __builtin_isfinite(rh[v]);
  return !finite || rho < gate || eint < emin;
}

// The update of one cell from its start-of-step state U0 and du = dtdx * (flux
// divergence): tracked density / energy / field increments, the drag propagator with its
// energy loss, the admissibility mark. o[8] is returned before the psi decay.
template <class T>
__device__ __forceinline__ void src_update(const T* U0, const T* du, T* o, const SrcParams<T>& S, int64_t idx,
                                           const StepParams<T>& P, bool ct) {
  T* const cr = S.corr;
  const int64_t n = S.stride;
  const bool trho = S.flags & SRC_TRACK_RHO;
  const bool te = S.flags & SRC_TRACK_E, tb = (S.flags & SRC_TRACK_B) && !ct;
  // The tracked slots are distinct arrays: all are read before any is written, so their
  // loads are in flight together (one memory latency, not one per read-after-write), and
  // the residuals are carried in registers and stored once at the end.
  T r4 = T(0), r8 = T(0), rb[3] = {T(0), T(0), T(0)};
  if (trho) r4 = cr[4 * n + idx];
  if (te) r8 = cr[8 * n + idx];
  if (tb) {
#pragma unroll
    for (int c = 0; c < 3; ++c) rb[c] = cr[(5 + c) * n + idx];
  }
  T rh[9];
#pragma unroll
  for (int v = 0; v < 4; ++v) rh[v] = U0[v] + du[v];
  rh[4] = comp_update(&r8, U0[4], du[4], te);
#pragma unroll
  for (int c = 0; c < 3; ++c) rh[5 + c] = comp_update(&rb[c], U0[5 + c], du[5 + c], tb);
  rh[8] = U0[8] + du[8];
  const bool bad = S.check != CHECK_NONE && src_inadmissible(rh, P);
#pragma unroll
  for (int v = 0; v < 9; ++v) o[v] = rh[v];
  if (S.flags & SRC_DRAG) {
    // glm_physics.jl:143-170,383-390
    // one reciprocal per density, shared by its divisions (glm_physics.jl divides; the
    // products differ from the quotients by about an ulp)
    const T rn = dmax(rh[0], P.smallr), ro = dmax(U0[0], P.smallr);
    const T iro = rcp(ro), irn = rcp(rn);
    const T vox = U0[1] * iro, voy = U0[2] * iro, voz = U0[3] * iro;
    const T dvx = rh[1] * irn - vox, dvy = rh[2] * irn - voy, dvz = rh[3] * irn - voz;
    const T wx = vox - S.vx0, wy = voy - S.vy0, wz = voz - S.vz0;
    const T vnx = S.vx0 + S.decay * wx + S.phi1 * dvx;
    const T vny = S.vy0 + S.decay * wy + S.phi1 * dvy;
    const T vnz = S.vz0 + S.decay * wz + S.phi1 * dvz;
    const T w2 = wx * wx + wy * wy + wz * wz;
    const T wdv = wx * dvx + wy * dvy + wz * dvz;
    const T dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
    const T iwx = (T(1) - S.decay) * wx + (T(1) - S.phi1) * dvx;
    const T iwy = (T(1) - S.decay) * wy + (T(1) - S.phi1) * dvy;
    const T iwz = (T(1) - S.decay) * wz + (T(1) - S.phi1) * dvz;
    const T qwork = rn * (S.A * w2 + S.C * wdv + S.D * dv2 + S.vx0 * iwx + S.vy0 * iwy + S.vz0 * iwz);
    const T nmx = rn * vnx, nmy = rn * vny, nmz = rn * vnz;
    const T ek = T(0.5) * (nmx * nmx + nmy * nmy + nmz * nmz) * irn;
    const T eb = T(0.5) * (rh[5] * rh[5] + rh[6] * rh[6] + rh[7] * rh[7]);
    const T emin = ek + eb + T(1e-30);
    o[0] = rn;
    o[1] = nmx;
    o[2] = nmy;
    o[3] = nmz;
    if (te) {
      T e = comp_update(&r8, rh[4], -qwork, true);
      if (e < emin) {
        r8 = T(0);
        e = emin;
      }
      o[4] = e;
    } else {
      o[4] = dmax(rh[4] - qwork, emin);
    }
  }
  if (bad) o[0] = -dmax(dabs(o[0]), dmax(P.smallr, P.llf_dmin));
  if (trho) cr[4 * n + idx] = r4 + du[0];
  if (te) cr[8 * n + idx] = r8;
  if (tb) {
#pragma unroll
    for (int c = 0; c < 3; ++c) cr[(5 + c) * n + idx] = rb[c];
  }
}

}  // namespace pmf
