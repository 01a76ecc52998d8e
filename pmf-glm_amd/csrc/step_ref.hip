// Per-cell GLM / CT step (the portable path, lib/MHDKernels/src/integrator_ref.jl:69-160):
// one work-item per cell rebuilds the Hancock prediction of itself and of its six face
// neighbours from a +-2 stencil and solves its six faces. Arbitrary per-axis boundary
// conditions (boundary.jl:17-43) are synthesised at load time. Also: the Riemann-flux
// batch entry used by unit tests and by the surrogate's training data.
#include "cellpredict.hpp"
#include "launch.hpp"
#include "sources.hpp"

namespace pmf {

// Both faces of one axis for the cell: D[v] = F_lo - F_hi (integrator_ref.jl:74-125).
// SRC: face states with the drag / radiation sources (integrator_ref.jl:28-48).
template <int D, int RECON, bool CT, bool SRC, class T>
__device__ void axis_faces(const Fields<const T>& U, T* __restrict__ ctp, const Grid& G, int i, int j, int k,
                           int64_t c, const T* m0c, const T* mhc, const T (*dLc)[9], const T (*dRc)[9],
                           const StepParams<T>& P, const SrcParams<T>& S, T* Dd) {
  constexpr int di = D == 0, dj = D == 1, dk = D == 2;
  T mh[9], dL[3][9], dR[3][9], m0[9];
  T Ls[9], Rs[9], Flo[9], Fhi[9];
  // lower face: left state from the lower neighbour's +D edge
  predict_cell<RECON>(U, G, i - di, j - dj, k - dk, P, mh, dL, dR, SRC ? m0 : nullptr);
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    Ls[v] = mh[v] + dR[D][v];
    Rs[v] = mhc[v] + dLc[D][v];
  }
  if constexpr (SRC) {
    src_face(Ls, m0, dR[D], S, periodic_index(G, i - di, j - dj, k - dk));
    src_face(Rs, m0c, dLc[D], S, c);
  }
  T bl = T(0), bh = T(0);
  int64_t cp = 0;
  if constexpr (CT) {
    int ii = i + di, jj = j + dj, kk = k + dk;
    ii = ii == G.nx ? 0 : ii;
    jj = jj == G.ny ? 0 : jj;
    kk = kk == G.nz ? 0 : kk;
    cp = ((int64_t)kk * G.ny + jj) * G.nx + ii;
    bl = ctp[D * G.nc + c];
    bh = ctp[D * G.nc + cp];
    Ls[5 + D] = bl;
    Rs[5 + D] = bl;
  }
  face_flux<D, CT>(Ls, Rs, Flo, P);
  // upper face: right state from the upper neighbour's -D edge
  predict_cell<RECON>(U, G, i + di, j + dj, k + dk, P, mh, dL, dR, SRC ? m0 : nullptr);
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    Ls[v] = mhc[v] + dRc[D][v];
    Rs[v] = mh[v] + dL[D][v];
  }
  if constexpr (SRC) {
    src_face(Ls, m0c, dRc[D], S, c);
    src_face(Rs, m0, dL[D], S, periodic_index(G, i + di, j + dj, k + dk));
  }
  if constexpr (CT) {
    Ls[5 + D] = bh;
    Rs[5 + D] = bh;
  }
  face_flux<D, CT>(Ls, Rs, Fhi, P);
#pragma unroll
  for (int v = 0; v < 9; ++v) Dd[v] = Flo[v] - Fhi[v];
  if constexpr (CT) {
    // integrator_ref.jl:129-132: lower-face mass flux and transverse-B fluxes
    const int64_t n = G.nc;
    ctp[(3 + D) * n + c] = Flo[0];
    if (D == 0) {
      ctp[6 * n + c] = Flo[6];
      ctp[7 * n + c] = Flo[7];
    } else if (D == 1) {
      ctp[8 * n + c] = Flo[5];
      ctp[9 * n + c] = Flo[7];
    } else {
      ctp[10 * n + c] = Flo[5];
      ctp[11 * n + c] = Flo[6];
    }
  }
}

template <int RECON, bool CT, bool SRC, class T>
__global__ __launch_bounds__(256) void k_step_ref(Fields<const T> U, Fields<T> O, T* __restrict__ ctp, Grid G,
                                                  StepParams<T> P, SrcParams<T> S) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G.nc) return;
  int i = (int)(c % G.nx);
  int j = (int)((c / G.nx) % G.ny);
  int k = (int)(c / ((int64_t)G.nx * G.ny));
  T mhc[9], dLc[3][9], dRc[3][9], m0c[9];
  predict_cell<RECON>(U, G, i, j, k, P, mhc, dLc, dRc, SRC ? m0c : nullptr);
  T Dx[9], Dy[9], Dz[9];
  axis_faces<0, RECON, CT, SRC>(U, ctp, G, i, j, k, c, m0c, mhc, dLc, dRc, P, S, Dx);
  axis_faces<1, RECON, CT, SRC>(U, ctp, G, i, j, k, c, m0c, mhc, dLc, dRc, P, S, Dy);
  axis_faces<2, RECON, CT, SRC>(U, ctp, G, i, j, k, c, m0c, mhc, dLc, dRc, P, S, Dz);
  const int64_t n = G.nc;
  T U0[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) U0[v] = U.v[v][c];
  if constexpr (CT) {
    // integrator_ref.jl:133-138: E = -v x B with the (source-modified) predicted velocity
    // and the start-of-step cell field
    T qc[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) qc[v] = mhc[v];
    if constexpr (SRC) {
      const T z[9] = {};
      src_face(qc, m0c, z, S, c);
    }
    ctp[12 * n + c] = qc[3] * U0[6] - qc[2] * U0[7];
    ctp[13 * n + c] = qc[1] * U0[7] - qc[3] * U0[5];
    ctp[14 * n + c] = qc[2] * U0[5] - qc[1] * U0[6];
  }
  // integrator_ref.jl:140-158
  if constexpr (SRC) {
    T du[9], o[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) du[v] = P.dtdx * ((Dx[v] + Dy[v]) + Dz[v]);
    src_update(U0, du, o, S, c, P, CT);
#pragma unroll
    for (int v = 0; v < 8; ++v) O.v[v][c] = o[v];
    O.v[8][c] = CT ? T(0) : o[8] * P.decay;
  } else {
#pragma unroll
    for (int v = 0; v < 8; ++v) O.v[v][c] = U0[v] + P.dtdx * ((Dx[v] + Dy[v]) + Dz[v]);
    O.v[8][c] = CT ? T(0) : (U0[8] + P.dtdx * ((Dx[8] + Dy[8]) + Dz[8])) * P.decay;
  }
}

template <class T>
void launch_step_ref(Fields<const T> U, Fields<T> O, T* ctp, const Grid& G, const StepParams<T>& P, bool ct,
                     hipStream_t st, const SrcParams<T>* S) {
  dim3 blk(256), grd((unsigned)((G.nc + 255) / 256));
  SrcParams<T> none{};
#define PMF_REF(R, C)                                                                                            \
  if (S) hipLaunchKernelGGL((k_step_ref<R, C, true, T>), grd, blk, 0, st, U, O, ctp, G, P, *S);                  \
  else hipLaunchKernelGGL((k_step_ref<R, C, false, T>), grd, blk, 0, st, U, O, ctp, G, P, none)
  if (ct) {
    if (P.recon == RECON_PPM) PMF_REF(RECON_PPM, true);
    else if (P.recon == RECON_PLM_MINMOD) PMF_REF(RECON_PLM_MINMOD, true);
    else PMF_REF(RECON_PLM, true);
  } else {
    if (P.recon == RECON_PPM) PMF_REF(RECON_PPM, false);
    else if (P.recon == RECON_PLM_MINMOD) PMF_REF(RECON_PLM_MINMOD, false);
    else PMF_REF(RECON_PLM, false);
  }
#undef PMF_REF
}

template <bool CT, class T>
__global__ __launch_bounds__(256) void k_riemann_batch(const T* __restrict__ L, const T* __restrict__ R,
                                                       T* __restrict__ F, int64_t n, int dir, StepParams<T> P) {
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  T l[9], r[9], o[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    l[v] = L[f * 9 + v];
    r[v] = R[f * 9 + v];
  }
  if (dir == 0) face_flux<0, CT>(l, r, o, P);
  else if (dir == 1) face_flux<1, CT>(l, r, o, P);
  else face_flux<2, CT>(l, r, o, P);
#pragma unroll
  for (int v = 0; v < 9; ++v) F[f * 9 + v] = o[v];
}

template <class T>
void launch_riemann_batch(const T* L, const T* R, T* F, int64_t n, int dir, bool ct, const StepParams<T>& P,
                          hipStream_t st) {
  if (n <= 0) return;
  dim3 blk(256), grd((unsigned)((n + 255) / 256));
  if (ct) hipLaunchKernelGGL((k_riemann_batch<true, T>), grd, blk, 0, st, L, R, F, n, dir, P);
  else hipLaunchKernelGGL((k_riemann_batch<false, T>), grd, blk, 0, st, L, R, F, n, dir, P);
}

template void launch_step_ref<float>(Fields<const float>, Fields<float>, float*, const Grid&,
                                     const StepParams<float>&, bool, hipStream_t, const SrcParams<float>*);
template void launch_step_ref<double>(Fields<const double>, Fields<double>, double*, const Grid&,
                                      const StepParams<double>&, bool, hipStream_t, const SrcParams<double>*);
template void launch_riemann_batch<float>(const float*, const float*, float*, int64_t, int, bool,
                                          const StepParams<float>&, hipStream_t);
template void launch_riemann_batch<double>(const double*, const double*, double*, int64_t, int, bool,
                                           const StepParams<double>&, hipStream_t);

}  // namespace pmf
