// Per-cell Hancock prediction from a +-2 stencil with boundary synthesis, shared by the
// per-cell step and the surrogate's face dump (integrator_ref.jl:17-26, boundary.jl:17-43).
#pragma once
#include "mhd_device.hpp"

namespace pmf {

// boundary.jl:17-31 (0-based): in-range source index and flip sign for one axis.
__device__ __forceinline__ int bc_map(int i, int n, int code, float& flip) {
  flip = 1.0f;
  if (code == BC_PERIODIC) {
    int r = i % n;
    return r < 0 ? r + n : r;
  }
  if (code == BC_OUTFLOW) return i < 0 ? 0 : (i >= n ? n - 1 : i);
  if (i < 0) {
    flip = -1.0f;
    return -1 - i;
  }
  if (i >= n) {
    flip = -1.0f;
    return 2 * n - 1 - i;
  }
  return i;
}

template <class T>
__device__ __forceinline__ void load_prim(const Fields<const T>& U, const Grid& G, int i, int j, int k, T* q,
                                          const StepParams<T>& P) {
  float fx, fy, fz;
  int ii = bc_map(i, G.nx, G.bc[0], fx), jj = bc_map(j, G.ny, G.bc[1], fy), kk = bc_map(k, G.nz, G.bc[2], fz);
  int64_t idx = ((int64_t)kk * G.ny + jj) * G.nx + ii;
  T c[9];
  c[0] = U.v[0][idx];
  c[1] = U.v[1][idx] * T(fx);
  c[2] = U.v[2][idx] * T(fy);
  c[3] = U.v[3][idx] * T(fz);
  c[4] = U.v[4][idx];
  c[5] = U.v[5][idx] * T(fx);
  c[6] = U.v[6][idx] * T(fy);
  c[7] = U.v[7][idx] * T(fz);
  c[8] = U.v[8][idx];
  cons2prim(c, q, P);
}

template <int RECON, class T>
__device__ __forceinline__ void offsets9(const T* qm, const T* q0, const T* qp, T* dL, T* dR) {
#pragma unroll
  for (int v = 0; v < 9; ++v) offsets<RECON>(qm[v], q0[v], qp[v], dL[v], dR[v]);
}

// Hancock prediction of logical cell (i,j,k) (integrator_ref.jl:17-26): mh and the
// per-axis face offsets.
template <int RECON, class T>
__device__ void predict_cell(const Fields<const T>& U, const Grid& G, int i, int j, int k, const StepParams<T>& P,
                             T* mh, T (*dL)[9], T (*dR)[9], T* m0out = nullptr) {
  T m0[9], qm[9], qp[9];
  load_prim(U, G, i, j, k, m0, P);
  load_prim(U, G, i - 1, j, k, qm, P);
  load_prim(U, G, i + 1, j, k, qp, P);
  offsets9<RECON>(qm, m0, qp, dL[0], dR[0]);
  load_prim(U, G, i, j - 1, k, qm, P);
  load_prim(U, G, i, j + 1, k, qp, P);
  offsets9<RECON>(qm, m0, qp, dL[1], dR[1]);
  load_prim(U, G, i, j, k - 1, qm, P);
  load_prim(U, G, i, j, k + 1, qp, P);
  offsets9<RECON>(qm, m0, qp, dL[2], dR[2]);
  // hancock_edges (glm_physics.jl:225-231)
  T u0[9];
  if constexpr (RECON != RECON_PPM) {
    hancock_plm(m0, dR[0], dR[1], dR[2], u0, P);
  } else {
    T acc[9], fp[9], fm[9], s[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) s[v] = m0[v] + dR[0][v];
    dflux<0>(s, fp, P.igm1);
#pragma unroll
    for (int v = 0; v < 9; ++v) s[v] = m0[v] + dL[0][v];
    dflux<0>(s, fm, P.igm1);
#pragma unroll
    for (int v = 0; v < 9; ++v) acc[v] = fp[v] - fm[v];
#pragma unroll
    for (int v = 0; v < 9; ++v) s[v] = m0[v] + dR[1][v];
    dflux<1>(s, fp, P.igm1);
#pragma unroll
    for (int v = 0; v < 9; ++v) s[v] = m0[v] + dL[1][v];
    dflux<1>(s, fm, P.igm1);
#pragma unroll
    for (int v = 0; v < 9; ++v) acc[v] += fp[v] - fm[v];
#pragma unroll
    for (int v = 0; v < 9; ++v) s[v] = m0[v] + dR[2][v];
    dflux<2>(s, fp, P.igm1);
#pragma unroll
    for (int v = 0; v < 9; ++v) s[v] = m0[v] + dL[2][v];
    dflux<2>(s, fm, P.igm1);
#pragma unroll
    for (int v = 0; v < 9; ++v) acc[v] += fp[v] - fm[v];
    prim2cons(m0, u0, P);
#pragma unroll
    for (int v = 0; v < 9; ++v) u0[v] -= P.hdtdx * acc[v];
  }
  cons2prim(u0, mh, P);
  if (m0out) {
#pragma unroll
    for (int v = 0; v < 9; ++v) m0out[v] = m0[v];
  }
}

// Periodic linear index of logical cell (i,j,k): the radiation correction slots are
// addressed this way whatever the boundary conditions (integrator_ref.jl:50-53).
__device__ __forceinline__ int64_t periodic_index(const Grid& G, int i, int j, int k) {
  int ii = i % G.nx, jj = j % G.ny, kk = k % G.nz;
  ii += ii < 0 ? G.nx : 0;
  jj += jj < 0 ? G.ny : 0;
  kk += kk < 0 ? G.nz : 0;
  return ((int64_t)kk * G.ny + jj) * G.nx + ii;
}

}  // namespace pmf
