// Device-side physics of the GLM-MHD MUSCL-Hancock step, for gfx950.
//
// Same scheme as lib/MHDKernels/src/glm_physics.jl (cited per function), organised for
// the GPU: fluxes are evaluated in the face-normal frame through compile-time component
// maps instead of permuted copies, reciprocals and reciprocal square roots are shared
// between the terms that divide by the same quantity, and the four-way HLLD flux choice
// is a select between the left and right star fans rather than four flux evaluations.
// Float kernels use the hardware v_rcp_f32 / v_sqrt_f32 / v_rsq_f32 (about 1 ulp);
// double kernels use exact division and square root.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmf {

enum { RECON_PLM = 0, RECON_PPM = 1, RECON_PLM_MINMOD = 2 };
enum { RS_HLLD = 0, RS_HLL = 1, RS_LLF = 2 };
enum { BC_PERIODIC = 0, BC_OUTFLOW = 1, BC_REFLECT = 2 };

// Scalars of one step, prepared on the host in the state's precision.
template <class T> struct StepParams {
  T gamma, gm1, igm1;        // gamma, gamma-1, 1/(gamma-1)
  T smallr, pfl, llf_dmin, llf_pmin;
  T ch, ich, ch2;            // GLM speed, 1/ch, ch^2 (ch = 0 for CT: ich unused)
  T dtdx, hdtdx, decay;      // dt/dx, dt/(2dx), psi factor
  int32_t recon, rsolve;
};

// The nine conserved fields, each a flat array of nc cells (state.jl:63-65: one array per
// variable, so a caller's separately allocated arrays are used in place).
template <class T> struct Fields {
  T* v[9];
};

struct Grid {
  int32_t nx, ny, nz;
  int32_t bc[3];
  int64_t nc;
};

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float dsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double dsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsq(double x) { return 1.0 / __builtin_sqrt(x); }
__device__ __forceinline__ float dmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double dmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ float dmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double dmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ float dabs(float a) { return __builtin_fabsf(a); }
__device__ __forceinline__ float dfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double dfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ double dabs(double a) { return __builtin_fabs(a); }

// cons2prim (glm_physics.jl:11-18): density floored before the velocity and kinetic
// energy, pressure floored at pfl.
// The fused step converts its LDS planes through several inlined copies of this
// function (column cells, halo cells, chunk start), so its arithmetic is pinned: explicit
// FMAs and no other contraction. A cell then gets the same primitives whichever copy
// converts it, and the step's result does not depend on where the tile seams fall.
// Returns 1/rho (rho = q[0]) for callers that need it again.
template <class T> __device__ __forceinline__ T cons2prim(const T* c, T* q, const StepParams<T>& P) {
#pragma clang fp contract(off)
  T rho = dmax(c[0], P.smallr);
  T ir = rcp(rho);
  T ek = T(0.5) * dfma(c[3], c[3], dfma(c[2], c[2], c[1] * c[1])) * ir;
  T em = T(0.5) * dfma(c[7], c[7], dfma(c[6], c[6], c[5] * c[5]));
  q[0] = rho;
  q[1] = c[1] * ir;
  q[2] = c[2] * ir;
  q[3] = c[3] * ir;
  q[4] = dmax(P.gm1 * (c[4] - ek - em), P.pfl);
  q[5] = c[5];
  q[6] = c[6];
  q[7] = c[7];
  q[8] = c[8];
  return ir;
}

// prim2cons (glm_physics.jl:19-23).
template <class T> __device__ __forceinline__ void prim2cons(const T* q, T* u, const StepParams<T>& P) {
  T rho = q[0];
  u[0] = rho;
  u[1] = rho * q[1];
  u[2] = rho * q[2];
  u[3] = rho * q[3];
  u[4] = q[4] * P.igm1 + T(0.5) * rho * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) +
         T(0.5) * (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]);
  u[5] = q[5];
  u[6] = q[6];
  u[7] = q[7];
  u[8] = q[8];
}

// fast_speed (glm_physics.jl:24-27) from density reciprocal ir, pressure p, |B|^2 b2 and
// normal field bn.
template <class T> __device__ __forceinline__ T cfast(T ir, T p, T b2, T bn, T gamma) {
  T c2 = gamma * p * ir;
  T d2 = T(0.5) * (b2 * ir + c2);
  return dsqrt(d2 + dsqrt(dmax(d2 * d2 - c2 * (bn * bn) * ir, T(0))));  // bn*bn: shared with the caller's
}

// Physical flux along axis D (glm_physics.jl:28-34 composed with the cyclic rotations
// of :36-45 and dir_flux :61): normal n = D, transverse a = D+1, b = D+2 (mod 3). The
// normal-B and psi slots are zero.
template <int D, class T> __device__ __forceinline__ void dflux(const T* q, T* f, T igm1) {
  constexpr int N = D, A = (D + 1) % 3, B = (D + 2) % 3;
  T rho = q[0], un = q[1 + N], ua = q[1 + A], ub = q[1 + B], p = q[4];
  T bn = q[5 + N], ba = q[5 + A], bb = q[5 + B];
  T b2 = bn * bn + ba * ba + bb * bb;
  T pt = p + T(0.5) * b2;
  T E = p * igm1 + T(0.5) * rho * (un * un + ua * ua + ub * ub) + T(0.5) * b2;
  T vb = un * bn + ua * ba + ub * bb;
  T m = rho * un;
  f[0] = m;
  f[1 + N] = m * un + pt - bn * bn;
  f[1 + A] = m * ua - bn * ba;
  f[1 + B] = m * ub - bn * bb;
  f[4] = (E + pt) * un - bn * vb;
  f[5 + N] = T(0);
  f[5 + A] = un * ba - ua * bn;
  f[5 + B] = un * bb - ub * bn;
  f[8] = T(0);
}

// Half the Hancock flux difference of a symmetric reconstruction along axis D:
// (dflux(q + d) - dflux(q - d)) / 2 (glm_physics.jl:228-230 with dL = -dR = -d, the PLM
// edges). Every flux term is a product of primitives; with x = X +- dX and y = Y +- dY
// the odd part of x*y is X dY + dX Y and its even part X Y + dX dY, so the difference
// is formed from those parts directly: equal in exact arithmetic, about 60 operations
// against the two fluxes' 75, and without the cancellation of subtracting two fluxes.
// The flux is  m = rho un;  m un + p + (ba^2 + bb^2 - bn^2)/2;  m ua - bn ba;
// m ub - bn bb;  g p un + m |v|^2/2 + (ba^2 + bb^2) un - bn (ua ba + ub bb) with
// g = gamma/(gamma-1) (the energy flux (E + pt) un - bn v.B expanded);  un ba - ua bn;
// un bb - ub bn. Contraction is pinned (explicit FMAs only) so every inlined copy gives
// the same bits.
// v2 = |v|^2 of q, the same in every direction (the caller forms it once).
template <int D, class T> __device__ __forceinline__ void dflux_odd(const T* q, const T* d, T v2, T* f, T igm1) {
#pragma clang fp contract(off)
  constexpr int N = D, A = (D + 1) % 3, B = (D + 2) % 3;
  const T rho = q[0], un = q[1 + N], ua = q[1 + A], ub = q[1 + B], p = q[4];
  const T bn = q[5 + N], ba = q[5 + A], bb = q[5 + B];
  const T dr = d[0], dun = d[1 + N], dua = d[1 + A], dub = d[1 + B], dp = d[4];
  const T dbn = d[5 + N], dba = d[5 + A], dbb = d[5 + B];
  const T me = dfma(rho, un, dr * dun);   // even part of rho un
  const T mo = dfma(rho, dun, dr * un);   // odd part
  const T t1 = dfma(ba, dba, bb * dbb);   // odd part of (ba^2 + bb^2) / 2
  f[0] = mo;
  f[1 + N] = dfma(me, dun, dfma(mo, un, dfma(-bn, dbn, dp + t1)));
  f[1 + A] = dfma(me, dua, dfma(mo, ua, -dfma(bn, dba, dbn * ba)));
  f[1 + B] = dfma(me, dub, dfma(mo, ub, -dfma(bn, dbb, dbn * bb)));
  const T ve = v2 + dfma(dun, dun, dfma(dua, dua, dub * dub));  // even |v|^2
  const T vo = dfma(un, dun, dfma(ua, dua, ub * dub));  // half the odd part of |v|^2
  const T qe = dfma(ba, ba, dfma(bb, bb, dfma(dba, dba, dbb * dbb)));  // even ba^2 + bb^2
  const T we = dfma(ua, ba, dfma(ub, bb, dfma(dua, dba, dub * dbb)));  // even ua ba + ub bb
  const T wo = dfma(ua, dba, dfma(dua, ba, dfma(ub, dbb, dub * bb)));  // odd part
  const T e4 = dfma(me, vo, (T(0.5) * mo) * ve);                      // odd m |v|^2 / 2
  const T q4 = dfma(qe, dun, (t1 + t1) * un);                         // odd (ba^2 + bb^2) un
  const T w4 = dfma(bn, wo, dbn * we);                                // odd bn (ua ba + ub bb)
  const T g4 = dfma(p, dun, dp * un);                                 // odd p un
  f[4] = dfma(igm1 + T(1), g4, (e4 + q4) - w4);
  f[5 + N] = T(0);
  f[5 + A] = dfma(un, dba, dfma(dun, ba, -dfma(ua, dbn, dua * bn)));
  f[5 + B] = dfma(un, dbb, dfma(dun, bb, -dfma(ub, dbn, dub * bn)));
  f[8] = T(0);
}

// MUSCL-Hancock half-step conserved state (glm_physics.jl:225-231, prim2cons :19-23) for
// PLM offsets (dL = -dR): prim2cons(m0) - (dt/dx) * the three half flux differences, the
// kinetic energy and dflux_odd sharing |v|^2. Pinned (explicit FMAs only): the fused and
// the per-cell steps call it, so they predict the same bits.
template <class T>
__device__ __forceinline__ void hancock_plm(const T* m0, const T* dRx, const T* dRy, const T* dRz, T* uh,
                                            const StepParams<T>& P) {
#pragma clang fp contract(off)
  const T v2 = dfma(m0[3], m0[3], dfma(m0[2], m0[2], m0[1] * m0[1]));
  T acc[9], f[9];
  dflux_odd<0>(m0, dRx, v2, acc, P.igm1);
  dflux_odd<1>(m0, dRy, v2, f, P.igm1);
#pragma unroll
  for (int v = 0; v < 9; ++v) acc[v] += f[v];
  dflux_odd<2>(m0, dRz, v2, f, P.igm1);
#pragma unroll
  for (int v = 0; v < 9; ++v) acc[v] += f[v];
  const T b2 = dfma(m0[7], m0[7], dfma(m0[6], m0[6], m0[5] * m0[5]));
  uh[0] = m0[0];
  uh[1] = m0[0] * m0[1];
  uh[2] = m0[0] * m0[2];
  uh[3] = m0[0] * m0[3];
  uh[4] = dfma(m0[4], P.igm1, dfma(T(0.5) * m0[0], v2, T(0.5) * b2));
#pragma unroll
  for (int v = 5; v < 9; ++v) uh[v] = m0[v];
#pragma unroll
  for (int v = 0; v < 9; ++v) uh[v] = dfma(-P.dtdx, acc[v], uh[v]);
}

// PLM face offsets (dL, dR) from the one-sided differences dl = q0 - qm and dr = qp - q0
// (glm_physics.jl:47-59), for offsets<RECON> and for callers that already hold a
// difference (the fused step carries each column's z difference from one plane to the
// next).
template <int RECON, class T> __device__ __forceinline__ void offsets_diff(T dl, T dr, T& dL, T& dR) {
  static_assert(RECON != RECON_PPM, "PLM only");
  if constexpr (sizeof(T) == 4) {
    // minmod(a, b) is the median of (a, b, 0), and MonCen the median of (2 minmod(a, b),
    // (a+b)/2, 0): one v_med3_f32 each instead of the sign/abs/min/compare chain (equal
    // to glm_physics.jl's form except where a*b underflows to zero)
    T mm = __builtin_amdgcn_fmed3f(dl, dr, T(0));
    T s;
    if constexpr (RECON == RECON_PLM_MINMOD) s = mm;
    else s = __builtin_amdgcn_fmed3f(T(2) * mm, T(0.5) * (dl + dr), T(0));
    s = T(0.5) * s;
    dL = -s;
    dR = s;
  } else {
    T s;
    if constexpr (RECON == RECON_PLM_MINMOD) {
      T m = dmin(dabs(dl), dabs(dr));
      s = dl >= T(0) ? m : -m;
    } else {
      T dc = T(0.5) * (dl + dr);
      T m = dmin(T(2) * dmin(dabs(dl), dabs(dr)), dabs(dc));
      s = dc >= T(0) ? m : -m;
    }
    s = dl * dr <= T(0) ? T(0) : T(0.5) * s;
    dL = -s;
    dR = s;
  }
}

// Limited face offsets (dL, dR) of one primitive from its two neighbours
// (glm_physics.jl:47-59 PLM MonCen / minmod, 188-200 PPM, 212-223).
template <int RECON, class T> __device__ __forceinline__ void offsets(T qm, T q0, T qp, T& dL, T& dR) {
  if constexpr (RECON == RECON_PPM) {
    // The unlimited parabola's edges are qL = q0 - b and qR = q0 + a. In these offsets the
    // CW84 monotonization of glm_physics.jl:192-200 reads: flat (both 0) when a*b <= 0;
    // otherwise diff > dq²/6 is |b| > 2|a| (then b := 2a), diff < -dq²/6 is |a| > 2|b|
    // (then a := 2b). For a and b of one sign that is a := minmod(a, 2b),
    // b := minmod(b, 2a), and minmod(x, y) = median(x, y, 0) covers the flat case too.
    // Every branch is continuous in (a, b), so rounding cannot flip a result by more than
    // rounding.
    T slope = (qp - qm) * T(0.25);
    T curve = (qm - T(2) * q0 + qp) * T(1.0 / 12.0);
    T a = slope + curve, b = slope - curve;
    if constexpr (sizeof(T) == 4) {
      dR = __builtin_amdgcn_fmed3f(a, T(2) * b, T(0));
      dL = -__builtin_amdgcn_fmed3f(b, T(2) * a, T(0));
    } else {
      const bool flat = a * b <= T(0);
      const T a2 = T(2) * a, b2 = T(2) * b;
      dR = flat ? T(0) : (dabs(a) > dabs(b2) ? b2 : a);
      dL = flat ? T(0) : -(dabs(b) > dabs(a2) ? a2 : b);
    }
  } else {
    offsets_diff<RECON>(q0 - qm, qp - q0, dL, dR);
  }
}

// ---------------------------------------------------------------------------
// Riemann solvers in the face-normal frame. States are primitive 9-vectors ordered
// (rho, un, ua, ub, p, bn, ba, bb, psi); fluxes come back in the same frame.

// Pinned forms (explicit FMAs, no other contraction) of the normal-frame flux, prim2cons
// and fast speed for the LLF and HLL solvers, so that every kernel that inlines them (the
// fused and the per-cell steps) gives the same bits.
template <class T> __device__ __forceinline__ void flux_n_pinned(const T* q, T* f, T igm1) {
#pragma clang fp contract(off)
  const T rho = q[0], un = q[1], ua = q[2], ub = q[3], p = q[4], bn = q[5], ba = q[6], bb = q[7];
  const T b2 = dfma(bb, bb, dfma(ba, ba, bn * bn));
  const T pt = dfma(T(0.5), b2, p);
  const T v2 = dfma(ub, ub, dfma(ua, ua, un * un));
  const T E = dfma(T(0.5), b2, dfma(T(0.5) * rho, v2, p * igm1));
  const T vb = dfma(ub, bb, dfma(ua, ba, un * bn));
  const T m = rho * un;
  f[0] = m;
  f[1] = dfma(-bn, bn, dfma(m, un, pt));
  f[2] = dfma(-bn, ba, m * ua);
  f[3] = dfma(-bn, bb, m * ub);
  f[4] = dfma(-bn, vb, (E + pt) * un);
  f[5] = T(0);
  f[6] = dfma(-ua, bn, un * ba);
  f[7] = dfma(-ub, bn, un * bb);
  f[8] = T(0);
}
template <class T> __device__ __forceinline__ void prim2cons_pinned(const T* q, T* u, T igm1) {
#pragma clang fp contract(off)
  const T rho = q[0];
  u[0] = rho;
  u[1] = rho * q[1];
  u[2] = rho * q[2];
  u[3] = rho * q[3];
  u[4] = dfma(T(0.5), dfma(q[7], q[7], dfma(q[6], q[6], q[5] * q[5])),
              dfma(T(0.5) * rho, dfma(q[3], q[3], dfma(q[2], q[2], q[1] * q[1])), q[4] * igm1));
  u[5] = q[5];
  u[6] = q[6];
  u[7] = q[7];
  u[8] = q[8];
}
template <class T> __device__ __forceinline__ T cfast_pinned(T ir, T p, T b2, T bn, T gamma) {
#pragma clang fp contract(off)
  const T c2 = (gamma * p) * ir;
  const T d2 = T(0.5) * dfma(b2, ir, c2);
  return dsqrt(d2 + dsqrt(dmax(dfma(d2, d2, -((c2 * (bn * bn)) * ir)), T(0))));
}

template <class T>
__device__ __forceinline__ void llf_n(const T* L, const T* R, T ch, T fbn, T fpsi, T* f, const StepParams<T>& P) {
  // glm_physics.jl:239-244
#pragma clang fp contract(off)
  T irL = rcp(L[0]), irR = rcp(R[0]);
  T b2L = dfma(L[7], L[7], dfma(L[6], L[6], L[5] * L[5]));
  T b2R = dfma(R[7], R[7], dfma(R[6], R[6], R[5] * R[5]));
  T smax = dmax(dmax(dabs(L[1]) + cfast_pinned(irL, L[4], b2L, L[5], P.gamma),
                     dabs(R[1]) + cfast_pinned(irR, R[4], b2R, R[5], P.gamma)), ch);
  T fL[9], fR[9], uL[9], uR[9];
  flux_n_pinned(L, fL, P.igm1);
  flux_n_pinned(R, fR, P.igm1);
  prim2cons_pinned(L, uL, P.igm1);
  prim2cons_pinned(R, uR, P.igm1);
  T hs = T(0.5) * smax;
#pragma unroll
  for (int i = 0; i < 9; ++i) f[i] = dfma(-hs, uR[i] - uL[i], T(0.5) * (fL[i] + fR[i]));
  f[5] = fbn;
  f[8] = fpsi;
}

template <class T>
__device__ __forceinline__ void hll_n(const T* L, const T* R, T bn, T fbn, T fpsi, T* f, const StepParams<T>& P) {
  // glm_physics.jl:245-252
#pragma clang fp contract(off)
  T irL = rcp(L[0]), irR = rcp(R[0]);
  T cfL = cfast_pinned(irL, L[4], dfma(L[7], L[7], dfma(L[6], L[6], L[5] * L[5])), bn, P.gamma);
  T cfR = cfast_pinned(irR, R[4], dfma(R[7], R[7], dfma(R[6], R[6], R[5] * R[5])), bn, P.gamma);
  T SL = dmin(dmin(L[1] - cfL, R[1] - cfR), T(0));
  T SR = dmax(dmax(L[1] + cfL, R[1] + cfR), T(0));
  T fL[9], fR[9], uL[9], uR[9];
  flux_n_pinned(L, fL, P.igm1);
  flux_n_pinned(R, fR, P.igm1);
  prim2cons_pinned(L, uL, P.igm1);
  prim2cons_pinned(R, uR, P.igm1);
  T inv = rcp(SR - SL);
  T sLR = SL * SR;
#pragma unroll
  for (int i = 0; i < 9; ++i) f[i] = dfma(sLR, uR[i] - uL[i], dfma(SR, fL[i], -(SL * fR[i]))) * inv;
  f[5] = fbn;
  f[8] = fpsi;
}

// HLLD (Miyoshi & Kusano 2005) as in glm_physics.jl:255-307. The flux is
//   F_side + S*(U*_side - U_side) + [inner]*S*_side*(U**_side - U*_side)
// with side = left when SM >= 0 and "inner" the two branches that use the double-star
// state; this equals the reference's four formulas branch by branch.
template <class T>
__device__ __forceinline__ void hlld_n(const T* L, const T* R, T bn, T fbn, T fpsi, T* f, const StepParams<T>& P) {
  const T dL = L[0], uL = L[1], vL = L[2], wL = L[3], pL = L[4], byL = L[6], bzL = L[7];
  const T dR = R[0], uR = R[1], vR = R[2], wR = R[3], pR = R[4], byR = R[6], bzR = R[7];
  const T bn2 = bn * bn;
  T irL = rcp(dL), irR = rcp(dR);
  T b2L = bn2 + byL * byL + bzL * bzL, b2R = bn2 + byR * byR + bzR * bzR;
  T ptL = pL + T(0.5) * b2L, ptR = pR + T(0.5) * b2R;
  T cfL = cfast(irL, pL, b2L, bn, P.gamma), cfR = cfast(irR, pR, b2R, bn, P.gamma);
  T cmax = dmax(cfL, cfR);
  T SL = dmin(dmin(uL, uR) - cmax, T(0));
  T SR = dmax(dmax(uL, uR) + cmax, T(0));
  T aL = (SL - uL) * dL, aR = (SR - uR) * dR;   // mass fluxes relative to the outer waves
  T iden = rcp(aR - aL);
  T SM = (aR * uR - aL * uL - ptR + ptL) * iden;
  T pts = (aR * ptL - aL * ptR + aL * aR * (uR - uL)) * iden;
  T iLM = rcp(SL - SM), iRM = rcp(SR - SM);
  T dLs = dmax(aL * iLM, T(1e-12)), dRs = dmax(aR * iRM, T(1e-12));
  T rqL = rsq(dLs), rqR = rsq(dRs);
  T sqL = dLs * rqL, sqR = dRs * rqR;
  T abn = dabs(bn);
  T SLs = SM - abn * rqL, SRs = SM + abn * rqR;
  // single-star transverse states (glm_physics.jl:255-263)
  T vyLs, vzLs, byLs, bzLs, vyRs, vzRs, byRs, bzRs;
  // a degenerate star state (|denK| < 1e-12) keeps the side's transverse velocity and
  // field: fac = 0 and gg = 1 give exactly that, with two selects per side instead of four
  {
    T denK = aL * (SL - SM) - bn2;
    bool deg = dabs(denK) < T(1e-12);
    T ik = rcp(denK);
    T fac = deg ? T(0) : bn * (SM - uL) * ik;
    T gg = deg ? T(1) : (aL * (SL - uL) - bn2) * ik;
    vyLs = vL - byL * fac;
    vzLs = wL - bzL * fac;
    byLs = byL * gg;
    bzLs = bzL * gg;
  }
  {
    T denK = aR * (SR - SM) - bn2;
    bool deg = dabs(denK) < T(1e-12);
    T ik = rcp(denK);
    T fac = deg ? T(0) : bn * (SM - uR) * ik;
    T gg = deg ? T(1) : (aR * (SR - uR) - bn2) * ik;
    vyRs = vR - byR * fac;
    vzRs = wR - bzR * fac;
    byRs = byR * gg;
    bzRs = bzR * gg;
  }
  // double-star state (glm_physics.jl:288-296): needs both single-star fans
  T sg = bn >= T(0) ? T(1) : T(-1);
  T isum = rcp(sqL + sqR);
  T vyss = (sqL * vyLs + sqR * vyRs + (byRs - byLs) * sg) * isum;
  T vzss = (sqL * vzLs + sqR * vzRs + (bzRs - bzLs) * sg) * isum;
  T byss = (sqL * byRs + sqR * byLs + sqL * sqR * (vyRs - vyLs) * sg) * isum;
  T bzss = (sqL * bzRs + sqR * bzLs + sqL * sqR * (vzRs - vzLs) * sg) * isum;
  T vbss = SM * bn + vyss * byss + vzss * bzss;
  // side selection (glm_physics.jl:297-305); energies and the star energies are needed
  // on the selected side only
  const bool left = SM >= T(0);
  const T d = left ? dL : dR, u = left ? uL : uR, v = left ? vL : vR, w = left ? wL : wR;
  const T p = left ? pL : pR, by = left ? byL : byR, bz = left ? bzL : bzR;
  const T pt = left ? ptL : ptR, b2 = left ? b2L : b2R;
  const T S = left ? SL : SR, Ss = left ? SLs : SRs, iM = left ? iLM : iRM;
  const T ds = left ? dLs : dRs, vys = left ? vyLs : vyRs, vzs = left ? vzLs : vzRs;
  const T bys = left ? byLs : byRs, bzs = left ? bzLs : bzRs;
  const T sq = left ? sqL : sqR;
  const T E = p * P.igm1 + T(0.5) * d * (u * u + v * v + w * w) + T(0.5) * b2;
  const T vb = u * bn + v * by + w * bz;
  const T vbs = SM * bn + vys * bys + vzs * bzs;
  const T Es = ((S - u) * E - pt * u + pts * SM + bn * (vb - vbs)) * iM;
  const T Ess = left ? Es - sq * (vbs - vbss) * sg : Es + sq * (vbs - vbss) * sg;
  const bool inner = left ? (SLs < T(0)) : (SRs >= T(0));
  const T a = inner ? Ss : T(0);
  // side physical flux (phys_flux_x of the state with normal field bn)
  T m = d * u;
  T F0 = m, F1 = m * u + pt - bn2, F2 = m * v - bn * by, F3 = m * w - bn * bz;
  T F4 = (E + pt) * u - bn * vb;
  T F6 = u * by - v * bn, F7 = u * bz - w * bn;
  f[0] = F0 + S * (ds - d);
  f[1] = F1 + S * (ds * SM - m);
  f[2] = F2 + S * (ds * vys - d * v) + a * (ds * (vyss - vys));
  f[3] = F3 + S * (ds * vzs - d * w) + a * (ds * (vzss - vzs));
  f[4] = F4 + S * (Es - E) + a * (Ess - Es);
  f[5] = fbn;
  f[6] = F6 + S * (bys - by) + a * (byss - bys);
  f[7] = F7 + S * (bzs - bz) + a * (bzss - bzs);
  f[8] = fpsi;
}

// riemann / riemann_ct (glm_physics.jl:317-348) on normal-frame states (modified in place:
// floors, and the normal field replaced by the GLM b_n*, or by the CT face value).
template <bool CT, class T, int RS = -1>
__device__ __forceinline__ void riemann_n(T* L, T* R, T* f, const StepParams<T>& P) {
  const int rsolve = RS >= 0 ? RS : P.rsolve;
  L[0] = dmax(L[0], P.smallr);
  L[4] = dmax(L[4], P.pfl);
  R[0] = dmax(R[0], P.smallr);
  R[4] = dmax(R[4], P.pfl);
  T bns, fbn, fpsi, ch;
  if constexpr (CT) {
    bns = L[5];
    fbn = T(0);
    fpsi = T(0);
    ch = T(0);
  } else {
    // glm_pair (glm_physics.jl:234-238), uncontracted: the psi flux ch^2 bn* ends in a
    // product, and a caller's flux difference F_lo - F_hi could otherwise fuse one face's
    // product into an FMA, so that equal fluxes would not cancel
#pragma clang fp contract(off)
    bns = T(0.5) * (L[5] + R[5]) - T(0.5) * (R[8] - L[8]) * P.ich;
    fbn = T(0.5) * (L[8] + R[8]) - T(0.5) * P.ch * (R[5] - L[5]);
    fpsi = P.ch2 * bns;
    ch = P.ch;
  }
  L[5] = bns;
  R[5] = bns;
  if (rsolve == RS_LLF) {
    llf_n(L, R, ch, fbn, fpsi, f, P);
  } else if (rsolve == RS_HLL) {
    hll_n(L, R, bns, fbn, fpsi, f, P);
  } else {
    bool uself = (P.llf_dmin > T(0) && dmin(L[0], R[0]) < P.llf_dmin) ||
                 (P.llf_pmin > T(0) && dmin(L[4], R[4]) < P.llf_pmin);
    if (uself)
      llf_n(L, R, ch, fbn, fpsi, f, P);
    else
      hlld_n(L, R, bns, fbn, fpsi, f, P);
  }
}

// Lab-frame face flux across an axis-D face from lab-frame primitive states.
template <int D, bool CT, class T, int RS = -1>
__device__ __forceinline__ void face_flux(const T* Lq, const T* Rq, T* F, const StepParams<T>& P) {
  constexpr int N = D, A = (D + 1) % 3, B = (D + 2) % 3;
  T L[9] = {Lq[0], Lq[1 + N], Lq[1 + A], Lq[1 + B], Lq[4], Lq[5 + N], Lq[5 + A], Lq[5 + B], Lq[8]};
  T R[9] = {Rq[0], Rq[1 + N], Rq[1 + A], Rq[1 + B], Rq[4], Rq[5 + N], Rq[5 + A], Rq[5 + B], Rq[8]};
  T f[9];
  riemann_n<CT, T, RS>(L, R, f, P);
  F[0] = f[0];
  F[1 + N] = f[1];
  F[1 + A] = f[2];
  F[1 + B] = f[3];
  F[4] = f[4];
  F[5 + N] = f[5];
  F[5 + A] = f[6];
  F[5 + B] = f[7];
  F[8] = f[8];
}

// GS05 upwind edge EMF (ct.jl:57-66): Ea*, Eb* the face EMFs on the a- and b-faces beside
// the edge, mA*, mB* the mass fluxes through them, Er** the cell EMFs of the four cells
// around it. Without contraction, so that the edge pass and the fused CT step agree bit for
// bit.
template <class T>
__device__ __forceinline__ T gs05(T Ea0, T Eam, T Eb0, T Ebm, T mA0, T mAm, T mB0, T mBm, T Er00, T Erm0, T Er0m,
                                  T Ermm) {
#pragma clang fp contract(off)
  auto up = [](T m, T lo, T hi) { return m > T(0) ? lo : (m < T(0) ? hi : T(0.5) * (lo + hi)); };
  T A1 = up(mA0, Erm0 - Ebm, Er00 - Eb0);
  T A3 = up(mAm, Ebm - Ermm, Eb0 - Er0m);
  T B1 = up(mB0, Er0m - Eam, Er00 - Ea0);
  T B3 = up(mBm, Eam - Ermm, Ea0 - Erm0);
  return T(0.25) * ((Ea0 + Eam + Eb0 + Ebm) + (A3 - A1) + (B3 - B1));
}

// Per-cell CFL speed max_d(|v_d| + c_fast,d) of a conserved state (timestep.jl:8-12).
template <class T> __device__ __forceinline__ T cell_wavespeed(const T* c, const StepParams<T>& P) {
  T q[9];
  const T ir = cons2prim(c, q, P);  // 1/q[0]
  T b2 = q[5] * q[5] + q[6] * q[6] + q[7] * q[7];
  T sx = dabs(q[1]) + cfast(ir, q[4], b2, q[5], P.gamma);
  T sy = dabs(q[2]) + cfast(ir, q[4], b2, q[6], P.gamma);
  T sz = dabs(q[3]) + cfast(ir, q[4], b2, q[7], P.gamma);
  return dmax(sx, dmax(sy, sz));
}

// NaN-propagating max (Julia's `maximum` returns NaN when any element is NaN).
template <class T> __device__ __forceinline__ T nanmax(T a, T b) {
  return (a != a || b != b) ? (a + b) : dmax(a, b);
}

}  // namespace pmf
