// Whole-grid reductions and the constrained-transport update kernels.
//
// Reductions run in two passes without atomics (so results are reproducible run to
// run): each block folds a grid-stride range with wave64 shuffles and LDS into one slot
// of a partials buffer, then a single block folds the partials. Sums are float64 with
// Neumaier compensation carried through every combine; maxima propagate NaN like the
// reference's `maximum`.
#include "launch.hpp"

namespace pmf {

constexpr int kBlk = 256;

struct KSum {
  double s, c;
};

__device__ __forceinline__ KSum ksum_add(KSum a, double x) {
  double t = a.s + x;
  double e = __builtin_fabs(a.s) >= __builtin_fabs(x) ? (a.s - t) + x : (x - t) + a.s;
  return KSum{t, a.c + e};
}
__device__ __forceinline__ KSum ksum_merge(KSum a, KSum b) {
  KSum r = ksum_add(a, b.s);
  r.c += b.c;
  return r;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = nanmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ KSum wave_sum(KSum v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    KSum w{__shfl_xor(v.s, o, 64), __shfl_xor(v.c, o, 64)};
    v = ksum_merge(v, w);
  }
  return v;
}

// Block-wide fold of NV compensated sums and NM maxima; thread 0 gets the result.
template <int NV, int NM>
__device__ void block_fold(KSum* sums, double* maxs) {
  __shared__ double sh[(kBlk / 64) * (2 * NV + NM) + 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < NV; ++v) sums[v] = wave_sum(sums[v]);
#pragma unroll
  for (int m = 0; m < NM; ++m) maxs[m] = wave_max(maxs[m]);
  constexpr int W = 2 * NV + NM;
  if (lane == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      sh[wave * W + 2 * v] = sums[v].s;
      sh[wave * W + 2 * v + 1] = sums[v].c;
    }
#pragma unroll
    for (int m = 0; m < NM; ++m) sh[wave * W + 2 * NV + m] = maxs[m];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlk / 64; ++w) {
#pragma unroll
      for (int v = 0; v < NV; ++v) sums[v] = ksum_merge(sums[v], KSum{sh[w * W + 2 * v], sh[w * W + 2 * v + 1]});
#pragma unroll
      for (int m = 0; m < NM; ++m) maxs[m] = nanmax(maxs[m], sh[w * W + 2 * NV + m]);
    }
  }
}

static inline unsigned red_grid(int64_t n) {
  int64_t b = (n + kBlk - 1) / kBlk;
  return (unsigned)(b < kRedBlocks ? (b < 1 ? 1 : b) : kRedBlocks);
}

// Final fold of nb partial slots of (NV sums, NM maxima) laid out as
// [s0 c0 s1 c1 ... m0 m1 ...] per block.
template <int NV, int NM>
__global__ __launch_bounds__(kBlk) void k_fold_final(const double* __restrict__ red, int nb, double* __restrict__ out) {
  KSum sums[NV > 0 ? NV : 1];
  double maxs[NM > 0 ? NM : 1];
  for (int v = 0; v < NV; ++v) sums[v] = KSum{0.0, 0.0};
  for (int m = 0; m < NM; ++m) maxs[m] = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlk) {
    const double* r = red + (int64_t)b * kRedSlots;
    for (int v = 0; v < NV; ++v) sums[v] = ksum_merge(sums[v], KSum{r[2 * v], r[2 * v + 1]});
    for (int m = 0; m < NM; ++m) maxs[m] = nanmax(maxs[m], r[2 * NV + m]);
  }
  block_fold<NV, NM>(sums, maxs);
  if (threadIdx.x == 0) {
    for (int v = 0; v < NV; ++v) out[v] = sums[v].s + sums[v].c;
    for (int m = 0; m < NM; ++m) out[NV + m] = maxs[m];
  }
}

template <int NV, int NM>
__device__ __forceinline__ void store_partial(double* red, KSum* sums, double* maxs) {
  if (threadIdx.x == 0) {
    double* r = red + (int64_t)blockIdx.x * kRedSlots;
    for (int v = 0; v < NV; ++v) {
      r[2 * v] = sums[v].s;
      r[2 * v + 1] = sums[v].c;
    }
    for (int m = 0; m < NM; ++m) r[2 * NV + m] = maxs[m];
  }
}

// ---- CFL wave speed (timestep.jl:4-23) ------------------------------------------
template <class T>
__global__ __launch_bounds__(kBlk) void k_wavespeed(Fields<const T> U, Grid G, StepParams<T> P,
                                                    double* __restrict__ red) {
  double m = 0.0;
  const int64_t n = G.nc;
  for (int64_t c = (int64_t)blockIdx.x * kBlk + threadIdx.x; c < n; c += (int64_t)gridDim.x * kBlk) {
    T q[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) q[v] = U.v[v][c];
    m = nanmax(m, (double)cell_wavespeed(q, P));
  }
  KSum none[1];
  block_fold<0, 1>(none, &m);
  store_partial<0, 1>(red, none, &m);
}

template <class T>
void launch_wavespeed(Fields<const T> U, const Grid& G, const StepParams<T>& P, double* red, double* out,
                      hipStream_t st) {
  unsigned nb = red_grid(G.nc);
  hipLaunchKernelGGL((k_wavespeed<T>), dim3(nb), dim3(kBlk), 0, st, U, G, P, red);
  hipLaunchKernelGGL((k_fold_final<0, 1>), dim3(1), dim3(kBlk), 0, st, red, (int)nb, out);
}

// ---- centred div B (diagnostics.jl:6-21): max and float64 sum of squares --------
__device__ __forceinline__ int bc_idx(int i, int n, int code, double& flip) {
  flip = 1.0;
  if (code == BC_PERIODIC) return i < 0 ? i + n : (i >= n ? i - n : i);
  if (code == BC_OUTFLOW) return i < 0 ? 0 : (i >= n ? n - 1 : i);
  if (i < 0) {
    flip = -1.0;
    return -1 - i;
  }
  if (i >= n) {
    flip = -1.0;
    return 2 * n - 1 - i;
  }
  return i;
}

template <class T>
__global__ __launch_bounds__(kBlk) void k_divb(Fields<const T> U, Grid G, T inv2dx, double* __restrict__ red) {
  const int64_t n = G.nc;
  const T *bx = U.v[5], *by = U.v[6], *bz = U.v[7];
  double m = 0.0;
  KSum s{0.0, 0.0};
  for (int64_t c = (int64_t)blockIdx.x * kBlk + threadIdx.x; c < n; c += (int64_t)gridDim.x * kBlk) {
    int i = (int)(c % G.nx), j = (int)((c / G.nx) % G.ny), k = (int)(c / ((int64_t)G.nx * G.ny));
    double fxp, fxm, fyp, fym, fzp, fzm;
    int ip = bc_idx(i + 1, G.nx, G.bc[0], fxp), im = bc_idx(i - 1, G.nx, G.bc[0], fxm);
    int jp = bc_idx(j + 1, G.ny, G.bc[1], fyp), jm = bc_idx(j - 1, G.ny, G.bc[1], fym);
    int kp = bc_idx(k + 1, G.nz, G.bc[2], fzp), km = bc_idx(k - 1, G.nz, G.bc[2], fzm);
    auto L = [&](int a, int b, int d) { return ((int64_t)d * G.ny + b) * G.nx + a; };
    T db = (bx[L(ip, j, k)] * T(fxp) - bx[L(im, j, k)] * T(fxm) + by[L(i, jp, k)] * T(fyp) -
            by[L(i, jm, k)] * T(fym) + bz[L(i, j, kp)] * T(fzp) - bz[L(i, j, km)] * T(fzm)) * inv2dx;
    double d = (double)dabs(db);
    m = nanmax(m, d);
    s = ksum_add(s, d * d);
  }
  block_fold<1, 1>(&s, &m);
  store_partial<1, 1>(red, &s, &m);
}

template <class T>
void launch_divb(Fields<const T> U, const Grid& G, T inv2dx, double* red, double* out, hipStream_t st) {
  unsigned nb = red_grid(G.nc);
  hipLaunchKernelGGL((k_divb<T>), dim3(nb), dim3(kBlk), 0, st, U, G, inv2dx, red);
  hipLaunchKernelGGL((k_fold_final<1, 1>), dim3(1), dim3(kBlk), 0, st, red, (int)nb, out);
}

// ---- conserved totals (diagnostics.jl:44-55) -------------------------------------
template <class T>
__global__ __launch_bounds__(kBlk) void k_totals(Fields<const T> U, Grid G, double* __restrict__ red) {
  const int64_t n = G.nc;
  KSum s[8];
#pragma unroll
  for (int v = 0; v < 8; ++v) s[v] = KSum{0.0, 0.0};
  for (int64_t c = (int64_t)blockIdx.x * kBlk + threadIdx.x; c < n; c += (int64_t)gridDim.x * kBlk) {
#pragma unroll
    for (int v = 0; v < 8; ++v) s[v] = ksum_add(s[v], (double)U.v[v][c]);
  }
  double none[1];
  block_fold<8, 0>(s, none);
  store_partial<8, 0>(red, s, none);
}

template <class T> void launch_totals(Fields<const T> U, const Grid& G, double* red, double* out8, hipStream_t st) {
  unsigned nb = red_grid(G.nc);
  hipLaunchKernelGGL((k_totals<T>), dim3(nb), dim3(kBlk), 0, st, U, G, red);
  hipLaunchKernelGGL((k_fold_final<8, 0>), dim3(1), dim3(kBlk), 0, st, red, (int)nb, out8);
}

// ---- constrained transport (ct.jl) -----------------------------------------------
struct Pc {
  int nx, ny, nz;
  __device__ __forceinline__ int64_t at(int i, int j, int k) const {
    i = i < 0 ? i + nx : (i >= nx ? i - nx : i);
    j = j < 0 ? j + ny : (j >= ny ? j - ny : j);
    k = k < 0 ? k + nz : (k >= nz ? k - nz : k);
    return ((int64_t)k * ny + j) * nx + i;
  }
};

// ct.jl:71-93: edge EMFs from the lower-face fluxes and the cell EMFs (gs05 in
// mhd_device.hpp, shared with the fused CT step's in-kernel edges).
template <class T> __global__ __launch_bounds__(kBlk) void k_ct_emf(T* __restrict__ ctp, Grid G) {
  const int64_t n = G.nc;
  int64_t c = (int64_t)blockIdx.x * kBlk + threadIdx.x;
  if (c >= n) return;
  Pc p{G.nx, G.ny, G.nz};
  int i = (int)(c % G.nx), j = (int)((c / G.nx) % G.ny), k = (int)(c / ((int64_t)G.nx * G.ny));
  const T *mfx = ctp + 3 * n, *mfy = ctp + 4 * n, *mfz = ctp + 5 * n;
  const T *fxby = ctp + 6 * n, *fxbz = ctp + 7 * n, *fybx = ctp + 8 * n, *fybz = ctp + 9 * n;
  const T *fzbx = ctp + 10 * n, *fzby = ctp + 11 * n;
  const T *ecx = ctp + 12 * n, *ecy = ctp + 13 * n, *ecz = ctp + 14 * n;
  int64_t cim = p.at(i - 1, j, k), cjm = p.at(i, j - 1, k), ckm = p.at(i, j, k - 1);
  int64_t cimjm = p.at(i - 1, j - 1, k), cjmkm = p.at(i, j - 1, k - 1), ckmim = p.at(i - 1, j, k - 1);
  ctp[17 * n + c] = gs05(-fxby[c], -fxby[cjm], fybx[c], fybx[cim], mfx[c], mfx[cjm], mfy[c], mfy[cim], ecz[c],
                         ecz[cim], ecz[cjm], ecz[cimjm]);
  ctp[15 * n + c] = gs05(-fybz[c], -fybz[ckm], fzby[c], fzby[cjm], mfy[c], mfy[ckm], mfz[c], mfz[cjm], ecx[c],
                         ecx[cjm], ecx[ckm], ecx[cjmkm]);
  ctp[16 * n + c] = gs05(-fzbx[c], -fzbx[cim], fxbz[c], fxbz[ckm], mfz[c], mfz[cim], mfx[c], mfx[ckm], ecy[c],
                         ecy[ckm], ecy[cim], ecy[ckmim]);
}

// ct.jl:96-129: face update by the discrete curl (optionally Kahan-compensated).
template <class T, bool COMP> __global__ __launch_bounds__(kBlk) void k_ct_faces(T* __restrict__ ctp, Grid G, T dtdx) {
  const int64_t n = G.nc;
  int64_t c = (int64_t)blockIdx.x * kBlk + threadIdx.x;
  if (c >= n) return;
  Pc p{G.nx, G.ny, G.nz};
  int i = (int)(c % G.nx), j = (int)((c / G.nx) % G.ny), k = (int)(c / ((int64_t)G.nx * G.ny));
  const T *ex = ctp + 15 * n, *ey = ctp + 16 * n, *ez = ctp + 17 * n;
  int64_t ip = p.at(i + 1, j, k), jp = p.at(i, j + 1, k), kp = p.at(i, j, k + 1);
  T inc[3] = {-(dtdx * ((ez[jp] - ez[c]) - (ey[kp] - ey[c]))), -(dtdx * ((ex[kp] - ex[c]) - (ez[ip] - ez[c]))),
              -(dtdx * ((ey[ip] - ey[c]) - (ex[jp] - ex[c])))};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    T b = ctp[d * n + c];
    if constexpr (COMP) {
      // glm_physics.jl:352-359
      T incr = inc[d] + ctp[(18 + d) * n + c];
      T upd = b + incr;
      ctp[(18 + d) * n + c] = incr - (upd - b);
      ctp[d * n + c] = upd;
    } else {
      ctp[d * n + c] = b + inc[d];
    }
  }
}

// ct.jl:132-143: cell B = face average.
template <class T> __global__ __launch_bounds__(kBlk) void k_ct_bcc(Fields<T> U, const T* __restrict__ ctp, Grid G) {
  const int64_t n = G.nc;
  int64_t c = (int64_t)blockIdx.x * kBlk + threadIdx.x;
  if (c >= n) return;
  Pc p{G.nx, G.ny, G.nz};
  int i = (int)(c % G.nx), j = (int)((c / G.nx) % G.ny), k = (int)(c / ((int64_t)G.nx * G.ny));
  U.v[5][c] = T(0.5) * (ctp[c] + ctp[p.at(i + 1, j, k)]);
  U.v[6][c] = T(0.5) * (ctp[n + c] + ctp[n + p.at(i, j + 1, k)]);
  U.v[7][c] = T(0.5) * (ctp[2 * n + c] + ctp[2 * n + p.at(i, j, k + 1)]);
}

// ct.jl:132-143 and timestep.jl:8-12 in one pass: the cell field from the faces, then
// the CFL speed of the finished cell (the step's final state), folded per block.
template <class T>
__global__ __launch_bounds__(kBlk) void k_ct_bcc_cfl(Fields<T> U, const T* __restrict__ ctp, Grid G, StepParams<T> P,
                                                     double* __restrict__ red) {
  const int64_t n = G.nc;
  Pc p{G.nx, G.ny, G.nz};
  double m = 0.0;
  for (int64_t c = (int64_t)blockIdx.x * kBlk + threadIdx.x; c < n; c += (int64_t)gridDim.x * kBlk) {
    int i = (int)(c % G.nx), j = (int)((c / G.nx) % G.ny), k = (int)(c / ((int64_t)G.nx * G.ny));
    T q[9];
#pragma unroll
    for (int v = 0; v < 5; ++v) q[v] = U.v[v][c];
    q[5] = T(0.5) * (ctp[c] + ctp[p.at(i + 1, j, k)]);
    q[6] = T(0.5) * (ctp[n + c] + ctp[n + p.at(i, j + 1, k)]);
    q[7] = T(0.5) * (ctp[2 * n + c] + ctp[2 * n + p.at(i, j, k + 1)]);
    q[8] = U.v[8][c];
    U.v[5][c] = q[5];
    U.v[6][c] = q[6];
    U.v[7][c] = q[7];
    m = nanmax(m, (double)cell_wavespeed(q, P));
  }
  KSum none[1];
  block_fold<0, 1>(none, &m);
  store_partial<0, 1>(red, none, &m);
}

template <class T>
void launch_ct_update(Fields<T> U, T* ctp, const Grid& G, T dtdx, bool compensate, hipStream_t st,
                      const StepParams<T>* P, double* red, double* dsmax, bool have_edges) {
  dim3 blk(kBlk), grd((unsigned)((G.nc + kBlk - 1) / kBlk));
  if (!have_edges) hipLaunchKernelGGL((k_ct_emf<T>), grd, blk, 0, st, ctp, G);
  if (compensate) hipLaunchKernelGGL((k_ct_faces<T, true>), grd, blk, 0, st, ctp, G, dtdx);
  else hipLaunchKernelGGL((k_ct_faces<T, false>), grd, blk, 0, st, ctp, G, dtdx);
  if (P && red && dsmax) {
    const unsigned nb = red_grid(G.nc);
    hipLaunchKernelGGL((k_ct_bcc_cfl<T>), dim3(nb), blk, 0, st, U, ctp, G, *P, red);
    hipLaunchKernelGGL((k_fold_final<0, 1>), dim3(1), blk, 0, st, red, (int)nb, dsmax);
  } else {
    hipLaunchKernelGGL((k_ct_bcc<T>), grd, blk, 0, st, U, ctp, G);
  }
}

// ct.jl:176-185: exact face divergence, maximum.
template <class T>
__global__ __launch_bounds__(kBlk) void k_divb_face(const T* __restrict__ ctp, Grid G, T invdx, double* __restrict__ red) {
  const int64_t n = G.nc;
  Pc p{G.nx, G.ny, G.nz};
  double m = 0.0;
  for (int64_t c = (int64_t)blockIdx.x * kBlk + threadIdx.x; c < n; c += (int64_t)gridDim.x * kBlk) {
    int i = (int)(c % G.nx), j = (int)((c / G.nx) % G.ny), k = (int)(c / ((int64_t)G.nx * G.ny));
    T d = ((ctp[p.at(i + 1, j, k)] - ctp[c]) + (ctp[n + p.at(i, j + 1, k)] - ctp[n + c]) +
           (ctp[2 * n + p.at(i, j, k + 1)] - ctp[2 * n + c])) * invdx;
    m = nanmax(m, (double)dabs(d));
  }
  KSum none[1];
  block_fold<0, 1>(none, &m);
  store_partial<0, 1>(red, none, &m);
}

template <class T> void launch_divb_face(const T* ctp, const Grid& G, T invdx, double* red, double* out, hipStream_t st) {
  unsigned nb = red_grid(G.nc);
  hipLaunchKernelGGL((k_divb_face<T>), dim3(nb), dim3(kBlk), 0, st, ctp, G, invdx, red);
  hipLaunchKernelGGL((k_fold_final<0, 1>), dim3(1), dim3(kBlk), 0, st, red, (int)nb, out);
}

// ct.jl:205-219: cell B <- face average, E keeps the thermal pressure, psi <- 0.
template <class T> __global__ __launch_bounds__(kBlk) void k_ct_install(Fields<T> U, const T* __restrict__ ctp, Grid G) {
  const int64_t n = G.nc;
  int64_t c = (int64_t)blockIdx.x * kBlk + threadIdx.x;
  if (c >= n) return;
  Pc p{G.nx, G.ny, G.nz};
  int i = (int)(c % G.nx), j = (int)((c / G.nx) % G.ny), k = (int)(c / ((int64_t)G.nx * G.ny));
  T bx = T(0.5) * (ctp[c] + ctp[p.at(i + 1, j, k)]);
  T by = T(0.5) * (ctp[n + c] + ctp[n + p.at(i, j + 1, k)]);
  T bz = T(0.5) * (ctp[2 * n + c] + ctp[2 * n + p.at(i, j, k + 1)]);
  T ox = U.v[5][c], oy = U.v[6][c], oz = U.v[7][c];
  U.v[4][c] += T(0.5) * ((bx * bx + by * by + bz * bz) - (ox * ox + oy * oy + oz * oz));
  U.v[5][c] = bx;
  U.v[6][c] = by;
  U.v[7][c] = bz;
  U.v[8][c] = T(0);
}

template <class T> __global__ __launch_bounds__(kBlk) void k_zero(T* __restrict__ a, int64_t n) {
  int64_t c = (int64_t)blockIdx.x * kBlk + threadIdx.x;
  if (c < n) a[c] = T(0);
}

template <class T> void launch_ct_install(Fields<T> U, T* ctp, const Grid& G, hipStream_t st) {
  dim3 blk(kBlk), grd((unsigned)((G.nc + kBlk - 1) / kBlk)), grz((unsigned)((3 * G.nc + kBlk - 1) / kBlk));
  hipLaunchKernelGGL((k_zero<T>), grz, blk, 0, st, ctp + 18 * G.nc, 3 * G.nc);
  hipLaunchKernelGGL((k_ct_install<T>), grd, blk, 0, st, U, ctp, G);
}

// ct.jl:310-320: face field = discrete curl of an edge-centred vector potential.
template <class T>
__global__ __launch_bounds__(kBlk) void k_ct_curl(const T* __restrict__ Ax, const T* __restrict__ Ay,
                                                  const T* __restrict__ Az, T* __restrict__ bx, T* __restrict__ by,
                                                  T* __restrict__ bz, Grid G, T invdx) {
  int64_t c = (int64_t)blockIdx.x * kBlk + threadIdx.x;
  if (c >= G.nc) return;
  Pc p{G.nx, G.ny, G.nz};
  int i = (int)(c % G.nx), j = (int)((c / G.nx) % G.ny), k = (int)(c / ((int64_t)G.nx * G.ny));
  bx[c] = ((Az[p.at(i, j + 1, k)] - Az[c]) - (Ay[p.at(i, j, k + 1)] - Ay[c])) * invdx;
  by[c] = ((Ax[p.at(i, j, k + 1)] - Ax[c]) - (Az[p.at(i + 1, j, k)] - Az[c])) * invdx;
  bz[c] = ((Ay[p.at(i + 1, j, k)] - Ay[c]) - (Ax[p.at(i, j + 1, k)] - Ax[c])) * invdx;
}

template <class T>
void launch_ct_curl(const T* Ax, const T* Ay, const T* Az, T* bx, T* by, T* bz, const Grid& G, T invdx,
                    hipStream_t st) {
  dim3 blk(kBlk), grd((unsigned)((G.nc + kBlk - 1) / kBlk));
  hipLaunchKernelGGL((k_ct_curl<T>), grd, blk, 0, st, Ax, Ay, Az, bx, by, bz, G, invdx);
}

#define PMF_INST(T)                                                                                             \
  template void launch_wavespeed<T>(Fields<const T>, const Grid&, const StepParams<T>&, double*, double*,       \
                                    hipStream_t);                                                               \
  template void launch_divb<T>(Fields<const T>, const Grid&, T, double*, double*, hipStream_t);                  \
  template void launch_totals<T>(Fields<const T>, const Grid&, double*, double*, hipStream_t);                  \
  template void launch_ct_update<T>(Fields<T>, T*, const Grid&, T, bool, hipStream_t, const StepParams<T>*, double*, \
                                     double*, bool);                                                               \
  template void launch_divb_face<T>(const T*, const Grid&, T, double*, double*, hipStream_t);                    \
  template void launch_ct_install<T>(Fields<T>, T*, const Grid&, hipStream_t);                                   \
  template void launch_ct_curl<T>(const T*, const T*, const T*, T*, T*, T*, const Grid&, T, hipStream_t);
PMF_INST(float)
PMF_INST(double)

}  // namespace pmf
