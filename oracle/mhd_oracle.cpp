// CPU oracle for the GLM-HLLD finite-volume step — TEST INFRASTRUCTURE ONLY.
//
// A restatement, in plain C++ (templated on double / float), of the reference's
// portable per-cell integrator and its helpers in yipihey/pmf-glm lib/MHDKernels:
//   glm_physics.jl:11-61,176-348   (cons2prim ... riemann_ct)
//   boundary.jl:17-56              (ghost-free BC synthesis)
//   integrator_ref.jl:17-27,69-160,210-220 (per-cell step, operation order kept)
//   timestep.jl:4-46               (CFL wave speed)
//   diagnostics.jl:6-55            (centred div B, conserved totals)
//   ct.jl:57-201,205-219,310-320   (GS05 CT edge EMF, face curl, cell B, face div)
// It keeps the reference's per-cell recompute structure (each cell rebuilds the
// Hancock states of itself and its six face neighbours and solves its six faces),
// so it is also the timed CPU baseline. It is compiled with -ffp-contract=off so the
// arithmetic is evaluated exactly in the order written.
//
// Only tests/, the package's smoke() and bench.py's cpu_baseline leg may load this
// library, and only as the checker. Nothing in the product path calls it.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <limits>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

template <class T> struct Q { T v[9]; };

enum { RECON_PLM = 0, RECON_PPM = 1, RECON_PLM_MINMOD = 2 };   // glm_physics.jl:202-204
enum { RS_HLLD = 0, RS_HLL = 1, RS_LLF = 2 };                    // glm_physics.jl:309-311
enum { BC_PERIODIC = 0, BC_OUTFLOW = 1, BC_REFLECT = 2 };        // boundary.jl:7-9

template <class T> struct Phys {
  T gamma, smallr, pfl, llf_dmin, llf_pmin;
  int recon, rsolve;
};

// glm_physics.jl:11-18
template <class T> inline Q<T> cons2prim(const Q<T>& c, T g, T smallr, T pfl) {
  T rho = std::max(c.v[0], smallr);
  T ir = T(1) / rho;
  T vx = c.v[1] * ir, vy = c.v[2] * ir, vz = c.v[3] * ir;
  T bx = c.v[5], by = c.v[6], bz = c.v[7];
  T ekin = T(0.5) * (c.v[1] * c.v[1] + c.v[2] * c.v[2] + c.v[3] * c.v[3]) * ir;
  T emag = T(0.5) * (bx * bx + by * by + bz * bz);
  T p = std::max((g - T(1)) * (c.v[4] - ekin - emag), pfl);
  return Q<T>{{rho, vx, vy, vz, p, bx, by, bz, c.v[8]}};
}

// glm_physics.jl:19-23
template <class T> inline Q<T> prim2cons(const Q<T>& q, T g) {
  T rho = q.v[0], vx = q.v[1], vy = q.v[2], vz = q.v[3], p = q.v[4];
  T bx = q.v[5], by = q.v[6], bz = q.v[7];
  T E = p / (g - T(1)) + T(0.5) * rho * (vx * vx + vy * vy + vz * vz) +
        T(0.5) * (bx * bx + by * by + bz * bz);
  return Q<T>{{rho, rho * vx, rho * vy, rho * vz, E, bx, by, bz, q.v[8]}};
}

// glm_physics.jl:24-27
template <class T> inline T fast_speed(const Q<T>& q, T g, T bn) {
  T rho = q.v[0];
  T c2 = g * q.v[4] / rho;
  T b2 = (q.v[5] * q.v[5] + q.v[6] * q.v[6] + q.v[7] * q.v[7]) / rho;
  T d2 = T(0.5) * (b2 + c2);
  return std::sqrt(d2 + std::sqrt(std::max(d2 * d2 - c2 * bn * bn / rho, T(0))));
}

// glm_physics.jl:28-34
template <class T> inline Q<T> phys_flux_x(const Q<T>& q, T g) {
  T rho = q.v[0], vx = q.v[1], vy = q.v[2], vz = q.v[3], p = q.v[4];
  T bx = q.v[5], by = q.v[6], bz = q.v[7];
  T b2 = bx * bx + by * by + bz * bz;
  T ptot = p + T(0.5) * b2;
  T E = p / (g - T(1)) + T(0.5) * rho * (vx * vx + vy * vy + vz * vz) + T(0.5) * b2;
  T vb = vx * bx + vy * by + vz * bz;
  return Q<T>{{rho * vx, rho * vx * vx + ptot - bx * bx, rho * vx * vy - bx * by,
               rho * vx * vz - bx * bz, (E + ptot) * vx - bx * vb, T(0),
               vx * by - vy * bx, vx * bz - vz * bx, T(0)}};
}

// glm_physics.jl:36-40 (dir 1,2,3 -> 0,1,2)
template <class T> inline Q<T> rot_to(const Q<T>& q, int dir) {
  if (dir == 0) return q;
  if (dir == 1) return Q<T>{{q.v[0], q.v[2], q.v[3], q.v[1], q.v[4], q.v[6], q.v[7], q.v[5], q.v[8]}};
  return Q<T>{{q.v[0], q.v[3], q.v[1], q.v[2], q.v[4], q.v[7], q.v[5], q.v[6], q.v[8]}};
}
// glm_physics.jl:41-45
template <class T> inline Q<T> rot_flux_from(const Q<T>& f, int dir) {
  if (dir == 0) return f;
  if (dir == 1) return Q<T>{{f.v[0], f.v[3], f.v[1], f.v[2], f.v[4], f.v[7], f.v[5], f.v[6], f.v[8]}};
  return Q<T>{{f.v[0], f.v[2], f.v[3], f.v[1], f.v[4], f.v[6], f.v[7], f.v[5], f.v[8]}};
}
// glm_physics.jl:61
template <class T> inline Q<T> dir_flux(const Q<T>& q, int dir, T g) {
  return rot_flux_from(phys_flux_x(rot_to(q, dir), g), dir);
}

// glm_physics.jl:47-51
template <class T> inline T moncen(T dl, T dr) {
  T dc = T(0.5) * (dl + dr);
  T sgn = dc >= T(0) ? T(1) : T(-1);
  T val = sgn * std::min(T(2) * std::min(std::fabs(dl), std::fabs(dr)), std::fabs(dc));
  return dl * dr <= T(0) ? T(0) : val;
}
// glm_physics.jl:54-57
template <class T> inline T minmod(T dl, T dr) {
  T sgn = dl >= T(0) ? T(1) : T(-1);
  return dl * dr <= T(0) ? T(0) : sgn * std::min(std::fabs(dl), std::fabs(dr));
}
// glm_physics.jl:190-197
template <class T> inline void ppm_monotonize(T qL, T qa, T qR, T& oL, T& oR) {
  T dq = qR - qL;
  T qmid = T(0.5) * (qL + qR);
  T dq2_6 = dq * dq / T(6);
  T diff = (qa - qmid) * dq;
  if ((qR - qa) * (qa - qL) <= T(0)) { oL = qa; oR = qa; return; }
  if (diff > dq2_6) { oL = T(3) * qa - T(2) * qR; oR = qR; return; }
  if (diff < -dq2_6) { oL = qL; oR = T(3) * qa - T(2) * qL; return; }
  oL = qL; oR = qR;
}
// glm_physics.jl:212-223: per-variable face offsets (dL, dR) from the cell mean.
template <class T>
inline void recon_offsets(const Q<T>& mm, const Q<T>& m0, const Q<T>& mp, int rec, Q<T>& dL, Q<T>& dR) {
  for (int i = 0; i < 9; ++i) {
    if (rec == RECON_PPM) {
      // glm_physics.jl:188-189 then 198-200
      T slope = (mp.v[i] - mm.v[i]) * T(0.25);
      T curve = (mm.v[i] - T(2) * m0.v[i] + mp.v[i]) / T(12);
      T qL = m0.v[i] - slope + curve, qR = m0.v[i] + slope + curve;
      T eL, eR;
      ppm_monotonize(qL, m0.v[i], qR, eL, eR);
      dL.v[i] = eL - m0.v[i];
      dR.v[i] = eR - m0.v[i];
    } else {
      T s = rec == RECON_PLM_MINMOD ? T(0.5) * minmod(m0.v[i] - mm.v[i], mp.v[i] - m0.v[i])
                                    : T(0.5) * moncen(m0.v[i] - mm.v[i], mp.v[i] - m0.v[i]);
      dL.v[i] = -s;
      dR.v[i] = s;
    }
  }
}

template <class T> inline Q<T> padd(const Q<T>& a, const Q<T>& d) {
  Q<T> r;
  for (int i = 0; i < 9; ++i) r.v[i] = a.v[i] + d.v[i];
  return r;
}

// glm_physics.jl:225-231
template <class T>
inline Q<T> hancock_edges(const Q<T>& m0, const Q<T>* dL, const Q<T>* dR, T dtdx, T g) {
  Q<T> u0 = prim2cons(m0, g);
  T h = T(0.5) * dtdx;
  Q<T> fxp = dir_flux(padd(m0, dR[0]), 0, g), fxm = dir_flux(padd(m0, dL[0]), 0, g);
  Q<T> fyp = dir_flux(padd(m0, dR[1]), 1, g), fym = dir_flux(padd(m0, dL[1]), 1, g);
  Q<T> fzp = dir_flux(padd(m0, dR[2]), 2, g), fzm = dir_flux(padd(m0, dL[2]), 2, g);
  Q<T> r;
  for (int i = 0; i < 9; ++i)
    r.v[i] = u0.v[i] - h * ((fxp.v[i] - fxm.v[i]) + (fyp.v[i] - fym.v[i]) + (fzp.v[i] - fzm.v[i]));
  return r;
}

// glm_physics.jl:234-238
template <class T> inline void glm_pair(T bnL, T bnR, T psiL, T psiR, T ch, T& bns, T& psis) {
  bns = T(0.5) * (bnL + bnR) - T(0.5) * (psiR - psiL) / ch;
  psis = T(0.5) * (psiL + psiR) - T(0.5) * ch * (bnR - bnL);
}

// glm_physics.jl:239-244
template <class T> inline Q<T> llf_x(const Q<T>& L, const Q<T>& R, T g, T ch, T fbn, T fpsi) {
  Q<T> fL = phys_flux_x(L, g), fR = phys_flux_x(R, g);
  Q<T> uL = prim2cons(L, g), uR = prim2cons(R, g);
  T smax = std::max(std::max(std::fabs(L.v[1]) + fast_speed(L, g, L.v[5]),
                             std::fabs(R.v[1]) + fast_speed(R, g, R.v[5])), ch);
  Q<T> f;
  for (int i = 0; i < 9; ++i) f.v[i] = T(0.5) * (fL.v[i] + fR.v[i]) - T(0.5) * smax * (uR.v[i] - uL.v[i]);
  f.v[5] = fbn; f.v[8] = fpsi;
  return f;
}
// glm_physics.jl:245-252
template <class T> inline Q<T> hll_x(const Q<T>& L, const Q<T>& R, T g, T ch, T bn, T fbn, T fpsi) {
  (void)ch;
  T cfL = fast_speed(L, g, bn), cfR = fast_speed(R, g, bn);
  T SL = std::min(std::min(L.v[1] - cfL, R.v[1] - cfR), T(0));
  T SR = std::max(std::max(L.v[1] + cfL, R.v[1] + cfR), T(0));
  Q<T> fL = phys_flux_x(L, g), fR = phys_flux_x(R, g);
  Q<T> uL = prim2cons(L, g), uR = prim2cons(R, g);
  T invd = T(1) / (SR - SL);
  Q<T> f;
  for (int i = 0; i < 9; ++i) f.v[i] = (SR * fL.v[i] - SL * fR.v[i] + SL * SR * (uR.v[i] - uL.v[i])) * invd;
  f.v[5] = fbn; f.v[8] = fpsi;
  return f;
}
// glm_physics.jl:255-263
template <class T>
inline void hlld_star(T d, T u, T v, T w, T by, T bz, T S, T bn, T SM, T& vs, T& ws, T& bys, T& bzs) {
  T denK = d * (S - u) * (S - SM) - bn * bn;
  if (std::fabs(denK) < T(1e-12)) { vs = v; ws = w; bys = by; bzs = bz; return; }
  T fac = bn * (SM - u) / denK;
  vs = v - by * fac;
  ws = w - bz * fac;
  bys = by * (d * (S - u) * (S - u) - bn * bn) / denK;
  bzs = bz * (d * (S - u) * (S - u) - bn * bn) / denK;
}
// glm_physics.jl:264-307
template <class T> inline Q<T> hlld_x(const Q<T>& L, const Q<T>& R, T g, T ch, T bn, T fbn, T fpsi) {
  (void)ch;
  T dL = L.v[0], uL = L.v[1], vL = L.v[2], wL = L.v[3], pL = L.v[4], byL = L.v[6], bzL = L.v[7];
  T dR = R.v[0], uR = R.v[1], vR = R.v[2], wR = R.v[3], pR = R.v[4], byR = R.v[6], bzR = R.v[7];
  T b2L = bn * bn + byL * byL + bzL * bzL, b2R = bn * bn + byR * byR + bzR * bzR;
  T ptL = pL + T(0.5) * b2L, ptR = pR + T(0.5) * b2R;
  T EL = pL / (g - T(1)) + T(0.5) * dL * (uL * uL + vL * vL + wL * wL) + T(0.5) * b2L;
  T ER = pR / (g - T(1)) + T(0.5) * dR * (uR * uR + vR * vR + wR * wR) + T(0.5) * b2R;
  T cfL = fast_speed(L, g, bn), cfR = fast_speed(R, g, bn);
  T SL = std::min(std::min(uL, uR) - std::max(cfL, cfR), T(0));
  T SR = std::max(std::max(uL, uR) + std::max(cfL, cfR), T(0));
  Q<T> UL{{dL, dL * uL, dL * vL, dL * wL, EL, bn, byL, bzL, T(0)}};
  Q<T> UR{{dR, dR * uR, dR * vR, dR * wR, ER, bn, byR, bzR, T(0)}};
  Q<T> FL = phys_flux_x(Q<T>{{dL, uL, vL, wL, pL, bn, byL, bzL, T(0)}}, g);
  Q<T> FR = phys_flux_x(Q<T>{{dR, uR, vR, wR, pR, bn, byR, bzR, T(0)}}, g);
  T den = (SR - uR) * dR - (SL - uL) * dL;
  T SM = ((SR - uR) * dR * uR - (SL - uL) * dL * uL - ptR + ptL) / den;
  T pts = ((SR - uR) * dR * ptL - (SL - uL) * dL * ptR + dL * dR * (SR - uR) * (SL - uL) * (uR - uL)) / den;
  T dLs = std::max(dL * (SL - uL) / (SL - SM), T(1e-12));
  T dRs = std::max(dR * (SR - uR) / (SR - SM), T(1e-12));
  T sqdLs = std::sqrt(dLs), sqdRs = std::sqrt(dRs);
  T SLs = SM - std::fabs(bn) / sqdLs, SRs = SM + std::fabs(bn) / sqdRs;
  T vyLs, vzLs, byLs, bzLs, vyRs, vzRs, byRs, bzRs;
  hlld_star(dL, uL, vL, wL, byL, bzL, SL, bn, SM, vyLs, vzLs, byLs, bzLs);
  hlld_star(dR, uR, vR, wR, byR, bzR, SR, bn, SM, vyRs, vzRs, byRs, bzRs);
  T vdotbL = uL * bn + vL * byL + wL * bzL, vdotbLs = SM * bn + vyLs * byLs + vzLs * bzLs;
  T ELs = ((SL - uL) * EL - ptL * uL + pts * SM + bn * (vdotbL - vdotbLs)) / (SL - SM);
  T vdotbR = uR * bn + vR * byR + wR * bzR, vdotbRs = SM * bn + vyRs * byRs + vzRs * bzRs;
  T ERs = ((SR - uR) * ER - ptR * uR + pts * SM + bn * (vdotbR - vdotbRs)) / (SR - SM);
  Q<T> ULs{{dLs, dLs * SM, dLs * vyLs, dLs * vzLs, ELs, bn, byLs, bzLs, T(0)}};
  Q<T> URs{{dRs, dRs * SM, dRs * vyRs, dRs * vzRs, ERs, bn, byRs, bzRs, T(0)}};
  T sgn = bn >= T(0) ? T(1) : T(-1);
  T invsum = T(1) / (sqdLs + sqdRs);
  T vyss = (sqdLs * vyLs + sqdRs * vyRs + (byRs - byLs) * sgn) * invsum;
  T vzss = (sqdLs * vzLs + sqdRs * vzRs + (bzRs - bzLs) * sgn) * invsum;
  T byss = (sqdLs * byRs + sqdRs * byLs + sqdLs * sqdRs * (vyRs - vyLs) * sgn) * invsum;
  T bzss = (sqdLs * bzRs + sqdRs * bzLs + sqdLs * sqdRs * (vzRs - vzLs) * sgn) * invsum;
  T vdotbss = SM * bn + vyss * byss + vzss * bzss;
  T ELss = ELs - sqdLs * (vdotbLs - vdotbss) * sgn;
  T ERss = ERs + sqdRs * (vdotbRs - vdotbss) * sgn;
  Q<T> ULss{{dLs, dLs * SM, dLs * vyss, dLs * vzss, ELss, bn, byss, bzss, T(0)}};
  Q<T> URss{{dRs, dRs * SM, dRs * vyss, dRs * vzss, ERss, bn, byss, bzss, T(0)}};
  Q<T> f;
  if (SLs >= T(0)) {
    for (int i = 0; i < 9; ++i) f.v[i] = FL.v[i] + SL * (ULs.v[i] - UL.v[i]);
  } else if (SM >= T(0)) {
    for (int i = 0; i < 9; ++i) f.v[i] = FL.v[i] + SLs * ULss.v[i] - (SLs - SL) * ULs.v[i] - SL * UL.v[i];
  } else if (SRs >= T(0)) {
    for (int i = 0; i < 9; ++i) f.v[i] = FR.v[i] + SRs * URss.v[i] - (SRs - SR) * URs.v[i] - SR * UR.v[i];
  } else {
    for (int i = 0; i < 9; ++i) f.v[i] = FR.v[i] + SR * (URs.v[i] - UR.v[i]);
  }
  f.v[5] = fbn; f.v[8] = fpsi;
  return f;
}

// glm_physics.jl:317-329
template <class T> inline Q<T> riemann(const Q<T>& Lq, const Q<T>& Rq, int dir, T ch, const Phys<T>& P) {
  Q<T> L = rot_to(Lq, dir), R = rot_to(Rq, dir);
  L.v[0] = std::max(L.v[0], P.smallr); L.v[4] = std::max(L.v[4], P.pfl);
  R.v[0] = std::max(R.v[0], P.smallr); R.v[4] = std::max(R.v[4], P.pfl);
  T bns, psis;
  glm_pair(L.v[5], R.v[5], L.v[8], R.v[8], ch, bns, psis);
  T fbn = psis, fpsi = ch * ch * bns;
  Q<T> Lc = L, Rc = R;
  Lc.v[5] = bns; Rc.v[5] = bns;
  bool uself = (P.llf_dmin > T(0) && std::min(L.v[0], R.v[0]) < P.llf_dmin) ||
               (P.llf_pmin > T(0) && std::min(L.v[4], R.v[4]) < P.llf_pmin);
  Q<T> f;
  if (P.rsolve == RS_LLF) f = llf_x(Lc, Rc, P.gamma, ch, fbn, fpsi);
  else if (P.rsolve == RS_HLL) f = hll_x(Lc, Rc, P.gamma, ch, bns, fbn, fpsi);
  else f = uself ? llf_x(Lc, Rc, P.gamma, ch, fbn, fpsi) : hlld_x(Lc, Rc, P.gamma, ch, bns, fbn, fpsi);
  return rot_flux_from(f, dir);
}

// glm_physics.jl:336-348 (constrained transport: no GLM pair, ch = 0)
template <class T> inline Q<T> riemann_ct(const Q<T>& Lq, const Q<T>& Rq, int dir, const Phys<T>& P) {
  Q<T> L0 = rot_to(Lq, dir), R0 = rot_to(Rq, dir);
  Q<T> L{{std::max(L0.v[0], P.smallr), L0.v[1], L0.v[2], L0.v[3], std::max(L0.v[4], P.pfl), L0.v[5], L0.v[6], L0.v[7], T(0)}};
  Q<T> R{{std::max(R0.v[0], P.smallr), R0.v[1], R0.v[2], R0.v[3], std::max(R0.v[4], P.pfl), L0.v[5], R0.v[6], R0.v[7], T(0)}};
  T bns = L.v[5], z = T(0);
  bool uself = (P.llf_dmin > z && std::min(L.v[0], R.v[0]) < P.llf_dmin) ||
               (P.llf_pmin > z && std::min(L.v[4], R.v[4]) < P.llf_pmin);
  Q<T> f;
  if (P.rsolve == RS_LLF) f = llf_x(L, R, P.gamma, z, z, z);
  else if (P.rsolve == RS_HLL) f = hll_x(L, R, P.gamma, z, bns, z, z);
  else f = uself ? llf_x(L, R, P.gamma, z, z, z) : hlld_x(L, R, P.gamma, z, bns, z, z);
  return rot_flux_from(f, dir);
}

// ---------------------------------------------------------------------------
// Grid access. Cells are 0-based (i,j,k); idx = (k*Ny + j)*Nx + i, i fastest
// (state.jl:1-13,95 with the 1-based offset removed).
struct Grid {
  int64_t nx, ny, nz;
  int bc[3];
};

// boundary.jl:17-31 (0-based): in-range source index and the flip sign.
inline int64_t bcidx(int64_t i, int64_t n, int code, int& flip) {
  flip = 1;
  if (code == BC_PERIODIC) return ((i % n) + n) % n;
  if (code == BC_OUTFLOW) return i < 0 ? 0 : (i >= n ? n - 1 : i);
  if (i < 0) { flip = -1; return -1 - i; }
  if (i >= n) { flip = -1; return 2 * n - 1 - i; }
  return i;
}

// boundary.jl:34-43 then cons2prim (boundary.jl:42-43 `_loadP_bc`)
template <class T>
inline Q<T> loadP(const T* const* u, const Grid& G, int64_t i, int64_t j, int64_t k, const Phys<T>& P) {
  int fx, fy, fz;
  int64_t ii = bcidx(i, G.nx, G.bc[0], fx), jj = bcidx(j, G.ny, G.bc[1], fy), kk = bcidx(k, G.nz, G.bc[2], fz);
  int64_t idx = (kk * G.ny + jj) * G.nx + ii;
  Q<T> c{{u[0][idx], u[1][idx] * T(fx), u[2][idx] * T(fy), u[3][idx] * T(fz), u[4][idx],
          u[5][idx] * T(fx), u[6][idx] * T(fy), u[7][idx] * T(fz), u[8][idx]}};
  return cons2prim(c, P.gamma, P.smallr, P.pfl);
}

template <class T> struct Edges {
  Q<T> m0, mh, dL[3], dR[3];
};

// integrator_ref.jl:17-26
template <class T>
inline Edges<T> mh_edges(const T* const* u, const Grid& G, int64_t i, int64_t j, int64_t k, T dtdx, const Phys<T>& P) {
  Edges<T> e;
  e.m0 = loadP(u, G, i, j, k, P);
  recon_offsets(loadP(u, G, i - 1, j, k, P), e.m0, loadP(u, G, i + 1, j, k, P), P.recon, e.dL[0], e.dR[0]);
  recon_offsets(loadP(u, G, i, j - 1, k, P), e.m0, loadP(u, G, i, j + 1, k, P), P.recon, e.dL[1], e.dR[1]);
  recon_offsets(loadP(u, G, i, j, k - 1, P), e.m0, loadP(u, G, i, j, k + 1, P), P.recon, e.dL[2], e.dR[2]);
  Q<T> uh = hancock_edges(e.m0, e.dL, e.dR, dtdx, P.gamma);
  e.mh = cons2prim(uh, P.gamma, P.smallr, P.pfl);
  return e;
}

inline int64_t pidx(const Grid& G, int64_t i, int64_t j, int64_t k) {
  int f;
  return (bcidx(k, G.nz, BC_PERIODIC, f) * G.ny + bcidx(j, G.ny, BC_PERIODIC, f)) * G.nx +
         bcidx(i, G.nx, BC_PERIODIC, f);
}

// ---------------------------------------------------------------------------
// Source terms of the drag / radiation-coupled steps (step_ref_drag!,
// step_ref_radiation!, step_cube_drag!, step_cube_radiation!).
template <class T> struct Src {
  bool drag = false, rad = false, track_rho = false, track_b = false, track_e = false;
  int check = 0;  // 0 none, 1 admissibility flag, 2 masked fallback (slot 4 > 0), 3 global fallback
  T q = 0;        // drag impulse Gamma dt
  T vx0 = 0, vy0 = 0, vz0 = 0;
  T* corr = nullptr;  // 9 slots of nc: psi, dv_x, dv_y, dv_z, rho perturbation (or mask), B_x, B_y, B_z, E residuals
};

// glm_physics.jl:63-65
template <class T> inline T drag_phi1(T q) {
  return std::fabs(q) < T(1e-4) ? T(1) - q / T(2) + q * q / T(6) : -std::expm1(-q) / q;
}

// glm_physics.jl:87-103: coefficients of Gamma * int_0^dt |w(t)|^2 dt
template <class T> inline void drag_work(T q, T& A, T& C, T& D) {
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

// glm_physics.jl:71-82 (drag_predict_face with the impulse, the target velocity v0)
template <class T>
inline Q<T> drag_predict_face(const Q<T>& q0, const Q<T>& qh, T impulse, T vx0, T vy0, T vz0) {
  T decay = std::exp(-impulse), phi1 = drag_phi1(impulse);
  Q<T> r = qh;
  r.v[1] = vx0 + decay * (q0.v[1] - vx0) + phi1 * (qh.v[1] - q0.v[1]);
  r.v[2] = vy0 + decay * (q0.v[2] - vy0) + phi1 * (qh.v[2] - q0.v[2]);
  r.v[3] = vz0 + decay * (q0.v[3] - vz0) + phi1 * (qh.v[3] - q0.v[3]);
  return r;
}

// integrator_ref.jl:33-39 (_facep_sources): drag-propagated face velocity (half-step
// impulse), plus the radiation velocity correction of the predicted cell.
template <class T>
inline Q<T> facep_sources(const Q<T>& m0, const Q<T>& mh, const Q<T>& d, const Src<T>& S, int64_t idx, int64_t nc) {
  Q<T> q = S.drag ? drag_predict_face(padd(m0, d), padd(mh, d), T(0.5) * S.q, S.vx0, S.vy0, S.vz0) : padd(mh, d);
  if (S.rad) {
    q.v[1] = q.v[1] + S.corr[nc + idx];
    q.v[2] = q.v[2] + S.corr[2 * nc + idx];
    q.v[3] = q.v[3] + S.corr[3 * nc + idx];
  }
  return q;
}

// glm_physics.jl:350-372: x + dx with a carried float residual (Kahan), optionally floored.
template <class T> inline T comp_update(T* res, T x, T dx, bool tracked) {
  if (!tracked) return x + dx;
  T carry = *res;
  T increment = dx + carry;
  T updated = x + increment;
  *res = increment - (updated - x);
  return updated;
}
template <class T> inline T comp_floor_update(T* res, T x, T dx, T floor, bool tracked) {
  if (!tracked) return std::max(x + dx, floor);
  T updated = comp_update(res, x, dx, true);
  if (updated < floor) {
    *res = T(0);
    return floor;
  }
  return updated;
}

// glm_physics.jl:143-170 and 383-390: the exact constant-force propagator applied to the
// conservative trial state; total energy loses the integrated drag work.
template <class T>
inline Q<T> drag_correct(const Q<T>& u0, const Q<T>& tr, T q, T vx0, T vy0, T vz0, T smallr, T smallE, T* eres,
                         bool tracked) {
  T decay = std::exp(-q), phi1 = drag_phi1(q), A, C, D;
  drag_work(q, A, C, D);
  T rn = std::max(tr.v[0], smallr), ro = std::max(u0.v[0], smallr);
  T vox = u0.v[1] / ro, voy = u0.v[2] / ro, voz = u0.v[3] / ro;
  T vtx = tr.v[1] / rn, vty = tr.v[2] / rn, vtz = tr.v[3] / rn;
  T dvx = vtx - vox, dvy = vty - voy, dvz = vtz - voz;
  T wx = vox - vx0, wy = voy - vy0, wz = voz - vz0;
  T vnx = vx0 + decay * wx + phi1 * dvx;
  T vny = vy0 + decay * wy + phi1 * dvy;
  T vnz = vz0 + decay * wz + phi1 * dvz;
  T w2 = wx * wx + wy * wy + wz * wz;
  T wdv = wx * dvx + wy * dvy + wz * dvz;
  T dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
  T iwx = (T(1) - decay) * wx + (T(1) - phi1) * dvx;
  T iwy = (T(1) - decay) * wy + (T(1) - phi1) * dvy;
  T iwz = (T(1) - decay) * wz + (T(1) - phi1) * dvz;
  T qwork = rn * (A * w2 + C * wdv + D * dv2 + vx0 * iwx + vy0 * iwy + vz0 * iwz);
  T nmx = rn * vnx, nmy = rn * vny, nmz = rn * vnz;
  T ek = T(0.5) * (nmx * nmx + nmy * nmy + nmz * nmz) / rn;
  T eb = T(0.5) * (tr.v[5] * tr.v[5] + tr.v[6] * tr.v[6] + tr.v[7] * tr.v[7]);
  T emin = ek + eb + smallE;
  Q<T> r = tr;
  r.v[0] = rn;
  r.v[1] = nmx;
  r.v[2] = nmy;
  r.v[3] = nmz;
  r.v[4] = comp_floor_update(eres, tr.v[4], -qwork, emin, tracked);
  return r;
}

// integrator_cube.jl:100-116
template <class T> inline bool inadmissible(const Q<T>& rh, T smallr, T pfl, T g, T llf_dmin) {
  T rho = rh.v[0];
  T gate = std::max(smallr, llf_dmin);
  T ir = T(1) / std::max(rho, smallr);
  T ek = T(0.5) * (rh.v[1] * rh.v[1] + rh.v[2] * rh.v[2] + rh.v[3] * rh.v[3]) * ir;
  T eb = T(0.5) * (rh.v[5] * rh.v[5] + rh.v[6] * rh.v[6] + rh.v[7] * rh.v[7]);
  T eint = rh.v[4] - ek - eb;
  T emin = pfl / (g - T(1));
  bool finite = std::isfinite(rho) && std::isfinite(eint);
  for (int v = 1; v < 9; ++v) finite = finite && std::isfinite(rh.v[v]);
  return !finite || rho < gate || eint < emin;
}

// integrator_ref.jl:69-160 for one cell (GLM, or CT when ctp != nullptr; no drag /
// radiation / tracking: the plain step of step_ref!, integrator_ref.jl:210-220).
//
// With sources (S != nullptr): the drag / radiation / tracking / admissibility terms of
// integrator_ref.jl:28-48,150-160 and, for `check`, of the cube kernel
// (integrator_cube.jl:68-116,330-340,423-441): a masked or global fallback zeroes the
// slopes of flagged cells and solves every face touching one with LLF.
template <class T>
inline bool src_masked(const Src<T>* S, const Grid& G, int64_t i, int64_t j, int64_t k) {
  if (!S || S->check < 2) return false;
  if (S->check == 3) return true;
  const int64_t nc = G.nx * G.ny * G.nz;
  return S->corr[4 * nc + pidx(G, i, j, k)] > T(0);
}

template <class T>
inline Edges<T> mh_edges_src(const T* const* u, const Grid& G, int64_t i, int64_t j, int64_t k, T dtdx,
                             const Phys<T>& P, const Src<T>* S) {
  if (!src_masked(S, G, i, j, k)) return mh_edges(u, G, i, j, k, dtdx, P);
  Edges<T> e;
  e.m0 = loadP(u, G, i, j, k, P);
  for (int d = 0; d < 3; ++d)
    for (int v = 0; v < 9; ++v) e.dL[d].v[v] = e.dR[d].v[v] = T(0);
  Q<T> uh = hancock_edges(e.m0, e.dL, e.dR, dtdx, P.gamma);
  e.mh = cons2prim(uh, P.gamma, P.smallr, P.pfl);
  return e;
}

template <class T>
void step_cell(const T* const* u, T* const* o, const Grid& G, int64_t c, T dtdx, T ch, T decay,
               const Phys<T>& P, T* ctp, const Src<T>* S = nullptr) {
  const int64_t nc = G.nx * G.ny * G.nz;
  int64_t i = c % G.nx, j = (c / G.nx) % G.ny, k = c / (G.nx * G.ny);
  Edges<T> ec = mh_edges_src(u, G, i, j, k, dtdx, P, S);
  Q<T> F[3][2];  // [dir][lo, hi]
  const int64_t di[3] = {1, 0, 0}, dj[3] = {0, 1, 0}, dk[3] = {0, 0, 1};
  const bool mc = src_masked(S, G, i, j, k);
  Phys<T> PL = P;
  PL.rsolve = RS_LLF;
  for (int d = 0; d < 3; ++d) {
    Edges<T> em = mh_edges_src(u, G, i - di[d], j - dj[d], k - dk[d], dtdx, P, S);
    Edges<T> ep = mh_edges_src(u, G, i + di[d], j + dj[d], k + dk[d], dtdx, P, S);
    Q<T> Ll, Rl, Lh, Rh;
    if (S) {
      const int64_t im = pidx(G, i - di[d], j - dj[d], k - dk[d]), ip = pidx(G, i + di[d], j + dj[d], k + dk[d]);
      Ll = facep_sources(em.m0, em.mh, em.dR[d], *S, im, nc);
      Rl = facep_sources(ec.m0, ec.mh, ec.dL[d], *S, c, nc);
      Lh = facep_sources(ec.m0, ec.mh, ec.dR[d], *S, c, nc);
      Rh = facep_sources(ep.m0, ep.mh, ep.dL[d], *S, ip, nc);
    } else {
      Ll = padd(em.mh, em.dR[d]);
      Rl = padd(ec.mh, ec.dL[d]);
      Lh = padd(ec.mh, ec.dR[d]);
      Rh = padd(ep.mh, ep.dL[d]);
    }
    if (S && S->check == 2) {
      const bool ml = src_masked(S, G, i - di[d], j - dj[d], k - dk[d]);
      const bool mp = src_masked(S, G, i + di[d], j + dj[d], k + dk[d]);
      F[d][0] = riemann(Ll, Rl, d, ch, (ml || mc) ? PL : P);
      F[d][1] = riemann(Lh, Rh, d, ch, (mc || mp) ? PL : P);
      continue;
    }
    if (ctp) {
      // integrator_ref.jl:83-87,100-104,117-121: the staggered face value is the
      // normal field of both states.
      T bl = ctp[d * nc + c], bh = ctp[d * nc + pidx(G, i + di[d], j + dj[d], k + dk[d])];
      Ll.v[5 + d] = bl; Rl.v[5 + d] = bl; Lh.v[5 + d] = bh; Rh.v[5 + d] = bh;
      F[d][0] = riemann_ct(Ll, Rl, d, P);
      F[d][1] = riemann_ct(Lh, Rh, d, P);
    } else {
      F[d][0] = riemann(Ll, Rl, d, ch, P);
      F[d][1] = riemann(Lh, Rh, d, ch, P);
    }
  }
  T U0[9];
  for (int v = 0; v < 9; ++v) U0[v] = u[v][c];
  if (ctp) {
    // integrator_ref.jl:127-139
    ctp[3 * nc + c] = F[0][0].v[0]; ctp[4 * nc + c] = F[1][0].v[0]; ctp[5 * nc + c] = F[2][0].v[0];
    ctp[6 * nc + c] = F[0][0].v[6]; ctp[7 * nc + c] = F[0][0].v[7];
    ctp[8 * nc + c] = F[1][0].v[5]; ctp[9 * nc + c] = F[1][0].v[7];
    ctp[10 * nc + c] = F[2][0].v[5]; ctp[11 * nc + c] = F[2][0].v[6];
    Q<T> zero{};
    const Q<T> qc = S ? facep_sources(ec.m0, ec.mh, zero, *S, c, nc) : ec.mh;
    ctp[12 * nc + c] = qc.v[3] * U0[6] - qc.v[2] * U0[7];
    ctp[13 * nc + c] = qc.v[1] * U0[7] - qc.v[3] * U0[5];
    ctp[14 * nc + c] = qc.v[2] * U0[5] - qc.v[1] * U0[6];
  }
  // integrator_ref.jl:140-158: U0 + dtdx*((Fxl-Fxh)+(Fyl-Fyh)+(Fzl-Fzh)), psi*decay
  T du[9];
  for (int v = 0; v < 9; ++v)
    du[v] = dtdx * ((F[0][0].v[v] - F[0][1].v[v]) + (F[1][0].v[v] - F[1][1].v[v]) + (F[2][0].v[v] - F[2][1].v[v]));
  if (!S) {
    for (int v = 0; v < 9; ++v) {
      T r = U0[v] + du[v];
      if (v == 8) r = ctp ? T(0) : r * decay;
      o[v][c] = r;
    }
    return;
  }
  // integrator_ref.jl:150-160 / integrator_cube.jl:423-441 with tracking and drag
  T* cr = S->corr;
  Q<T> rh;
  if (S->track_rho) cr[4 * nc + c] = cr[4 * nc + c] + du[0];
  rh.v[0] = U0[0] + du[0];
  rh.v[1] = U0[1] + du[1];
  rh.v[2] = U0[2] + du[2];
  rh.v[3] = U0[3] + du[3];
  rh.v[4] = comp_update(cr ? &cr[8 * nc + c] : nullptr, U0[4], du[4], S->track_e);
  const bool tb = S->track_b && !ctp;
  rh.v[5] = comp_update(cr ? &cr[5 * nc + c] : nullptr, U0[5], du[5], tb);
  rh.v[6] = comp_update(cr ? &cr[6 * nc + c] : nullptr, U0[6], du[6], tb);
  rh.v[7] = comp_update(cr ? &cr[7 * nc + c] : nullptr, U0[7], du[7], tb);
  rh.v[8] = U0[8] + du[8];
  const bool bad = S->check != 0 && inadmissible(rh, P.smallr, P.pfl, P.gamma, P.llf_dmin);
  Q<T> U0q;
  for (int v = 0; v < 9; ++v) U0q.v[v] = U0[v];
  Q<T> r = S->drag ? drag_correct(U0q, rh, S->q, S->vx0, S->vy0, S->vz0, P.smallr, T(1e-30),
                                  cr ? &cr[8 * nc + c] : nullptr, S->track_e)
                   : rh;
  for (int v = 0; v < 8; ++v) o[v][c] = r.v[v];
  if (bad) o[0][c] = -std::max(std::fabs(r.v[0]), std::max(P.smallr, P.llf_dmin));
  o[8][c] = ctp ? T(0) : r.v[8] * decay;
}

template <class T>
void step_ref(const T* U, T* O, const Grid& G, T dtdx, T ch, T decay, const Phys<T>& P, T* ctp,
              const Src<T>* S = nullptr) {
  const int64_t nc = G.nx * G.ny * G.nz;
  const T* u[9];
  T* o[9];
  for (int v = 0; v < 9; ++v) { u[v] = U + v * nc; o[v] = O + v * nc; }
  if (S && S->rad) u[8] = S->corr;  // psi enters through slot 0 of the correction buffer
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < nc; ++c) step_cell(u, o, G, c, dtdx, ch, decay, P, ctp, S);
}

// timestep.jl:4-15: per-cell max_d(|v_d| + c_f(q, gamma, B_d)), then the max.
template <class T> T max_wavespeed(const T* U, const Grid& G, const Phys<T>& P) {
  const int64_t nc = G.nx * G.ny * G.nz;
  T smax = T(0);
  bool nan = false;
#pragma omp parallel for reduction(max : smax) reduction(|| : nan)
  for (int64_t c = 0; c < nc; ++c) {
    Q<T> cq;
    for (int v = 0; v < 9; ++v) cq.v[v] = U[v * nc + c];
    Q<T> q = cons2prim(cq, P.gamma, P.smallr, P.pfl);
    T sx = std::fabs(q.v[1]) + fast_speed(q, P.gamma, q.v[5]);
    T sy = std::fabs(q.v[2]) + fast_speed(q, P.gamma, q.v[6]);
    T sz = std::fabs(q.v[3]) + fast_speed(q, P.gamma, q.v[7]);
    T w = std::max(sx, std::max(sy, sz));
    if (std::isnan(w)) nan = true;
    smax = std::max(smax, w);
  }
  return nan ? std::numeric_limits<T>::quiet_NaN() : smax;
}

// diagnostics.jl:6-21: BC-aware centred |div B|, max over cells.
template <class T> T max_divb(const T* U, const Grid& G, T inv2dx) {
  const int64_t nc = G.nx * G.ny * G.nz;
  const T *bx = U + 5 * nc, *by = U + 6 * nc, *bz = U + 7 * nc;
  T dmax = T(0);
#pragma omp parallel for reduction(max : dmax)
  for (int64_t c = 0; c < nc; ++c) {
    int64_t i = c % G.nx, j = (c / G.nx) % G.ny, k = c / (G.nx * G.ny);
    int fxp, fxm, fyp, fym, fzp, fzm;
    int64_t ip = bcidx(i + 1, G.nx, G.bc[0], fxp), im = bcidx(i - 1, G.nx, G.bc[0], fxm);
    int64_t jp = bcidx(j + 1, G.ny, G.bc[1], fyp), jm = bcidx(j - 1, G.ny, G.bc[1], fym);
    int64_t kp = bcidx(k + 1, G.nz, G.bc[2], fzp), km = bcidx(k - 1, G.nz, G.bc[2], fzm);
    auto L = [&](int64_t a, int64_t b, int64_t d) { return (d * G.ny + b) * G.nx + a; };
    T db = (bx[L(ip, j, k)] * T(fxp) - bx[L(im, j, k)] * T(fxm) + by[L(i, jp, k)] * T(fyp) -
            by[L(i, jm, k)] * T(fym) + bz[L(i, j, kp)] * T(fzp) - bz[L(i, j, km)] * T(fzm)) * inv2dx;
    dmax = std::max(dmax, std::fabs(db));
  }
  return dmax;
}

// ct.jl:57-66
template <class T>
inline T ct_gs05(T Ea0, T Eam, T Eb0, T Ebm, T mA0, T mAm, T mB0, T mBm, T Er00, T Erm0, T Er0m, T Ermm) {
  const T h = T(0.5);
  auto up = [h](T m, T lower, T upper) { return m > T(0) ? lower : (m < T(0) ? upper : h * (lower + upper)); };
  T A1 = up(mA0, Erm0 - Ebm, Er00 - Eb0);
  T A3 = up(mAm, Ebm - Ermm, Eb0 - Er0m);
  T B1 = up(mB0, Er0m - Eam, Er00 - Ea0);
  T B3 = up(mBm, Eam - Ermm, Ea0 - Erm0);
  return T(0.25) * ((Ea0 + Eam + Eb0 + Ebm) + (A3 - A1) + (B3 - B1));
}

// ct.jl:158-173 (ct_update!): edge EMFs (71-93), face curl plain (96-107) or Kahan
// (118-129), then cell B = face average (132-143). ctp holds the 21 CT slots.
template <class T> void ct_update(T* U, T* ctp, const Grid& G, T dtdx, bool compensate) {
  const int64_t nc = G.nx * G.ny * G.nz;
  T *bxf = ctp, *byf = ctp + nc, *bzf = ctp + 2 * nc;
  const T *mfx = ctp + 3 * nc, *mfy = ctp + 4 * nc, *mfz = ctp + 5 * nc;
  const T *fxby = ctp + 6 * nc, *fxbz = ctp + 7 * nc, *fybx = ctp + 8 * nc, *fybz = ctp + 9 * nc;
  const T *fzbx = ctp + 10 * nc, *fzby = ctp + 11 * nc;
  const T *ecx = ctp + 12 * nc, *ecy = ctp + 13 * nc, *ecz = ctp + 14 * nc;
  T *ex = ctp + 15 * nc, *ey = ctp + 16 * nc, *ez = ctp + 17 * nc;
  T *rbx = ctp + 18 * nc, *rby = ctp + 19 * nc, *rbz = ctp + 20 * nc;
#pragma omp parallel for
  for (int64_t c = 0; c < nc; ++c) {
    int64_t i = c % G.nx, j = (c / G.nx) % G.ny, k = c / (G.nx * G.ny);
    int64_t cim = pidx(G, i - 1, j, k), cjm = pidx(G, i, j - 1, k), ckm = pidx(G, i, j, k - 1);
    int64_t cimjm = pidx(G, i - 1, j - 1, k), cjmkm = pidx(G, i, j - 1, k - 1), ckmim = pidx(G, i - 1, j, k - 1);
    ez[c] = ct_gs05(-fxby[c], -fxby[cjm], fybx[c], fybx[cim], mfx[c], mfx[cjm], mfy[c], mfy[cim],
                    ecz[c], ecz[cim], ecz[cjm], ecz[cimjm]);
    ex[c] = ct_gs05(-fybz[c], -fybz[ckm], fzby[c], fzby[cjm], mfy[c], mfy[ckm], mfz[c], mfz[cjm],
                    ecx[c], ecx[cjm], ecx[ckm], ecx[cjmkm]);
    ey[c] = ct_gs05(-fzbx[c], -fzbx[cim], fxbz[c], fxbz[ckm], mfz[c], mfz[cim], mfx[c], mfx[ckm],
                    ecy[c], ecy[ckm], ecy[cim], ecy[ckmim]);
  }
#pragma omp parallel for
  for (int64_t c = 0; c < nc; ++c) {
    int64_t i = c % G.nx, j = (c / G.nx) % G.ny, k = c / (G.nx * G.ny);
    int64_t ip = pidx(G, i + 1, j, k), jp = pidx(G, i, j + 1, k), kp = pidx(G, i, j, k + 1);
    T incx = -(dtdx * ((ez[jp] - ez[c]) - (ey[kp] - ey[c])));
    T incy = -(dtdx * ((ex[kp] - ex[c]) - (ez[ip] - ez[c])));
    T incz = -(dtdx * ((ey[ip] - ey[c]) - (ex[jp] - ex[c])));
    if (compensate) {
      // glm_physics.jl:352-359 (_compensated_update!)
      T* bf[3] = {bxf, byf, bzf};
      T* rb[3] = {rbx, rby, rbz};
      T inc[3] = {incx, incy, incz};
      for (int d = 0; d < 3; ++d) {
        T increment = inc[d] + rb[d][c];
        T updated = bf[d][c] + increment;
        rb[d][c] = increment - (updated - bf[d][c]);
        bf[d][c] = updated;
      }
    } else {
      bxf[c] -= dtdx * ((ez[jp] - ez[c]) - (ey[kp] - ey[c]));
      byf[c] -= dtdx * ((ex[kp] - ex[c]) - (ez[ip] - ez[c]));
      bzf[c] -= dtdx * ((ey[ip] - ey[c]) - (ex[jp] - ex[c]));
    }
  }
#pragma omp parallel for
  for (int64_t c = 0; c < nc; ++c) {
    int64_t i = c % G.nx, j = (c / G.nx) % G.ny, k = c / (G.nx * G.ny);
    U[5 * nc + c] = T(0.5) * (bxf[c] + bxf[pidx(G, i + 1, j, k)]);
    U[6 * nc + c] = T(0.5) * (byf[c] + byf[pidx(G, i, j + 1, k)]);
    U[7 * nc + c] = T(0.5) * (bzf[c] + bzf[pidx(G, i, j, k + 1)]);
  }
}

// ct.jl:176-185: exact face divergence, max over cells.
template <class T> T max_divb_face(const T* ctp, const Grid& G, T invdx) {
  const int64_t nc = G.nx * G.ny * G.nz;
  const T *bxf = ctp, *byf = ctp + nc, *bzf = ctp + 2 * nc;
  T dmax = T(0);
#pragma omp parallel for reduction(max : dmax)
  for (int64_t c = 0; c < nc; ++c) {
    int64_t i = c % G.nx, j = (c / G.nx) % G.ny, k = c / (G.nx * G.ny);
    T d = std::fabs(((bxf[pidx(G, i + 1, j, k)] - bxf[c]) + (byf[pidx(G, i, j + 1, k)] - byf[c]) +
                     (bzf[pidx(G, i, j, k + 1)] - bzf[c])) * invdx);
    dmax = std::max(dmax, d);
  }
  return dmax;
}

// ct.jl:205-219 (_ct_cell_from_faces_k!) after copying the faces in and zeroing the
// remainders (ct.jl:221-236).
template <class T> void ct_install_faces(T* U, T* ctp, const Grid& G, const T* bx, const T* by, const T* bz) {
  const int64_t nc = G.nx * G.ny * G.nz;
  std::memcpy(ctp, bx, nc * sizeof(T));
  std::memcpy(ctp + nc, by, nc * sizeof(T));
  std::memcpy(ctp + 2 * nc, bz, nc * sizeof(T));
  std::memset(ctp + 18 * nc, 0, 3 * nc * sizeof(T));
  const T *bxf = ctp, *byf = ctp + nc, *bzf = ctp + 2 * nc;
#pragma omp parallel for
  for (int64_t c = 0; c < nc; ++c) {
    int64_t i = c % G.nx, j = (c / G.nx) % G.ny, k = c / (G.nx * G.ny);
    T bxc = T(0.5) * (bxf[c] + bxf[pidx(G, i + 1, j, k)]);
    T byc = T(0.5) * (byf[c] + byf[pidx(G, i, j + 1, k)]);
    T bzc = T(0.5) * (bzf[c] + bzf[pidx(G, i, j, k + 1)]);
    T ox = U[5 * nc + c], oy = U[6 * nc + c], oz = U[7 * nc + c];
    U[4 * nc + c] += T(0.5) * ((bxc * bxc + byc * byc + bzc * bzc) - (ox * ox + oy * oy + oz * oz));
    U[5 * nc + c] = bxc; U[6 * nc + c] = byc; U[7 * nc + c] = bzc; U[8 * nc + c] = T(0);
  }
}

// ct.jl:310-320: face field = discrete curl of an edge-centred vector potential.
template <class T> void ct_curl(const T* Ax, const T* Ay, const T* Az, T* bx, T* by, T* bz, const Grid& G, T invdx) {
  const int64_t nc = G.nx * G.ny * G.nz;
#pragma omp parallel for
  for (int64_t c = 0; c < nc; ++c) {
    int64_t i = c % G.nx, j = (c / G.nx) % G.ny, k = c / (G.nx * G.ny);
    bx[c] = ((Az[pidx(G, i, j + 1, k)] - Az[c]) - (Ay[pidx(G, i, j, k + 1)] - Ay[c])) * invdx;
    by[c] = ((Ax[pidx(G, i, j, k + 1)] - Ax[c]) - (Az[pidx(G, i + 1, j, k)] - Az[c])) * invdx;
    bz[c] = ((Ay[pidx(G, i + 1, j, k)] - Ay[c]) - (Ax[pidx(G, i, j + 1, k)] - Ax[c])) * invdx;
  }
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C ABI for the Python test harness (ctypes). `U`/`O` are 9*nc contiguous SoA
// arrays in conserved order (rho, mx, my, mz, E, Bx, By, Bz, psi).
using namespace orc;

template <class T>
static Phys<T> mkphys(double g, double smallr, double pfl, double dmin, double pmin, int recon, int rs) {
  return Phys<T>{T(g), T(smallr), T(pfl), T(dmin), T(pmin), recon, rs};
}

#define ORC_DEFINE(SUF, T)                                                                              \
  extern "C" void orc_step_##SUF(const T* U, T* O, int64_t nx, int64_t ny, int64_t nz, const int* bc,   \
                                 int recon, int rs, double dtdx, double ch, double decay, double g,     \
                                 double smallr, double pfl, double dmin, double pmin, T* ctp) {         \
    Grid G{nx, ny, nz, {bc[0], bc[1], bc[2]}};                                                          \
    step_ref<T>(U, O, G, T(dtdx), T(ch), T(decay), mkphys<T>(g, smallr, pfl, dmin, pmin, recon, rs),    \
                ctp);                                                                                   \
  }                                                                                                     \
  /* flags: 1 drag, 2 radiation, 4/8/16 track rho/B/E, check mode in bits 8-9 */                          \
  extern "C" void orc_step_src_##SUF(const T* U, T* O, int64_t nx, int64_t ny, int64_t nz, const int* bc,  \
                                     int recon, int rs, double dtdx, double ch, double decay, double g,    \
                                     double smallr, double pfl, double dmin, double pmin, T* ctp,           \
                                     int flags, double q, double vx0, double vy0, double vz0, T* corr) {    \
    Grid G{nx, ny, nz, {bc[0], bc[1], bc[2]}};                                                           \
    Src<T> S;                                                                                            \
    S.drag = flags & 1; S.rad = flags & 2; S.track_rho = flags & 4; S.track_b = flags & 8;               \
    S.track_e = flags & 16; S.check = (flags >> 8) & 3;                                                  \
    S.q = T(q); S.vx0 = T(vx0); S.vy0 = T(vy0); S.vz0 = T(vz0); S.corr = corr;                           \
    step_ref<T>(U, O, G, T(dtdx), T(ch), T(decay), mkphys<T>(g, smallr, pfl, dmin, pmin, recon, rs), ctp, \
                &S);                                                                                     \
  }                                                                                                      \
  extern "C" double orc_max_wavespeed_##SUF(const T* U, int64_t nx, int64_t ny, int64_t nz, double g,   \
                                            double smallr, double pfl) {                                \
    Grid G{nx, ny, nz, {0, 0, 0}};                                                                      \
    return double(max_wavespeed<T>(U, G, mkphys<T>(g, smallr, pfl, 0, 0, 0, 0)));                       \
  }                                                                                                     \
  extern "C" double orc_max_divb_##SUF(const T* U, int64_t nx, int64_t ny, int64_t nz, const int* bc,   \
                                       double dx) {                                                     \
    Grid G{nx, ny, nz, {bc[0], bc[1], bc[2]}};                                                          \
    return double(max_divb<T>(U, G, T(1) / (T(2) * T(dx))));                                            \
  }                                                                                                     \
  extern "C" void orc_ct_update_##SUF(T* U, T* ctp, int64_t nx, int64_t ny, int64_t nz, double dtdx,    \
                                      int compensate) {                                                 \
    Grid G{nx, ny, nz, {0, 0, 0}};                                                                      \
    ct_update<T>(U, ctp, G, T(dtdx), compensate != 0);                                                  \
  }                                                                                                     \
  extern "C" double orc_max_divb_face_##SUF(const T* ctp, int64_t nx, int64_t ny, int64_t nz,           \
                                            double dx) {                                                \
    Grid G{nx, ny, nz, {0, 0, 0}};                                                                      \
    return double(max_divb_face<T>(ctp, G, T(1) / T(dx)));                                              \
  }                                                                                                     \
  extern "C" void orc_ct_install_faces_##SUF(T* U, T* ctp, int64_t nx, int64_t ny, int64_t nz,          \
                                             const T* bx, const T* by, const T* bz) {                   \
    Grid G{nx, ny, nz, {0, 0, 0}};                                                                      \
    ct_install_faces<T>(U, ctp, G, bx, by, bz);                                                         \
  }                                                                                                     \
  extern "C" void orc_ct_curl_##SUF(const T* Ax, const T* Ay, const T* Az, T* bx, T* by, T* bz,         \
                                    int64_t nx, int64_t ny, int64_t nz, double dx) {                    \
    Grid G{nx, ny, nz, {0, 0, 0}};                                                                      \
    ct_curl<T>(Ax, Ay, Az, bx, by, bz, G, T(1) / T(dx));                                                \
  }                                                                                                     \
  /* One Riemann flux per face: L, R, F are n x 9 row-major primitive states / fluxes. */              \
  extern "C" void orc_riemann_batch_##SUF(const T* L, const T* R, T* F, int64_t n, int dir, int ct,     \
                                          double ch, double g, double smallr, double pfl, double dmin,  \
                                          double pmin, int rs) {                                        \
    Phys<T> P = mkphys<T>(g, smallr, pfl, dmin, pmin, 0, rs);                                           \
    for (int64_t f = 0; f < n; ++f) {                                                                   \
      Q<T> l, r;                                                                                        \
      for (int v = 0; v < 9; ++v) { l.v[v] = L[f * 9 + v]; r.v[v] = R[f * 9 + v]; }                     \
      Q<T> o = ct ? riemann_ct(l, r, dir, P) : riemann(l, r, dir, T(ch), P);                            \
      for (int v = 0; v < 9; ++v) F[f * 9 + v] = o.v[v];                                                \
    }                                                                                                   \
  }                                                                                                     \
  /* Hancock predictor of one cell from its primitive mean and 3 neighbour pairs. */                   \
  extern "C" void orc_predict_batch_##SUF(const T* M, T* MH, T* DL, T* DR, int64_t n, int recon,        \
                                          double dtdx, double g, double smallr, double pfl) {           \
    /* M: n x 7 x 9 primitives (self, -x, +x, -y, +y, -z, +z); outputs mh (n x 9), dL/dR (n x 3 x 9) */ \
    for (int64_t f = 0; f < n; ++f) {                                                                   \
      Q<T> q[7];                                                                                        \
      for (int s = 0; s < 7; ++s)                                                                       \
        for (int v = 0; v < 9; ++v) q[s].v[v] = M[(f * 7 + s) * 9 + v];                                \
      Q<T> dl[3], dr[3];                                                                                \
      for (int d = 0; d < 3; ++d) recon_offsets(q[1 + 2 * d], q[0], q[2 + 2 * d], recon, dl[d], dr[d]);  \
      Q<T> uh = hancock_edges(q[0], dl, dr, T(dtdx), T(g));                                             \
      Q<T> mh = cons2prim(uh, T(g), T(smallr), T(pfl));                                                 \
      for (int v = 0; v < 9; ++v) {                                                                     \
        MH[f * 9 + v] = mh.v[v];                                                                        \
        for (int d = 0; d < 3; ++d) {                                                                   \
          DL[(f * 3 + d) * 9 + v] = dl[d].v[v];                                                         \
          DR[(f * 3 + d) * 9 + v] = dr[d].v[v];                                                         \
        }                                                                                               \
      }                                                                                                 \
    }                                                                                                   \
  }

ORC_DEFINE(f64, double)
ORC_DEFINE(f32, float)

// Drag propagator coefficients at impulse q: exp(-q), phi1(q), A, C, D (glm_physics.jl:63-103).
extern "C" void orc_drag_coeffs_f64(double q, double* out) {
  out[0] = std::exp(-q);
  out[1] = drag_phi1(q);
  drag_work(q, out[2], out[3], out[4]);
}
extern "C" void orc_drag_coeffs_f32(double q, double* out) {
  float A, C, D, qf = float(q);
  out[0] = std::exp(-qf);
  out[1] = drag_phi1(qf);
  drag_work(qf, A, C, D);
  out[2] = A;
  out[3] = C;
  out[4] = D;
}

extern "C" int orc_num_threads(void) {
#ifdef _OPENMP
  int n = 1;
#pragma omp parallel
  {
#pragma omp single
    n = omp_get_num_threads();
  }
  return n;
#else
  return 1;
#endif
}
