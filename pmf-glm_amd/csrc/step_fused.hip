// Fused, z-marching GLM step for periodic grids (the throughput path; the MI355X
// counterpart of lib/MHDKernels/src/integrator_cube.jl:295-476, same unsplit scheme).
//
// Work decomposition. A workgroup owns a 16x16 column tile in x-y and marches it
// through a chunk of z planes. Its 320 threads (5 wave64s) are bound for the whole
// march to one column each: 256 owned columns plus the 64 columns of the one-cell halo
// ring (corners excluded), which need their Hancock prediction for the tile's boundary
// faces. Every thread keeps the primitive states of its column for planes m-1, m, m+1
// in registers (the z stencil); the in-plane neighbours come from one LDS plane that
// also holds the second halo ring and the four ring corners. Per plane m:
//   1. convert the next plane of the own column (and, for a few threads, of the extra
//      halo cells) from conserved to primitive;
//   2. Hancock-predict plane m in every column (cons2prim of the half-step state);
//   3. solve each x and y face of the tile once (L state through LDS, R in registers)
//      and the z face m-1/2 of each owned column (both states in registers);
//   4. update plane m-1 of the owned columns with the reference's flux grouping, damp
//      psi, store, and fold the new cell's CFL speed into a running maximum.
// Every cell is predicted once per tile pass (1.25x with the ring), every face is solved
// once, and the global state is read once and written once per step. Workgroups are
// remapped so that consecutive tiles land on the same XCD (and its L2), which then
// serves the halo re-reads.
#include "launch.hpp"

namespace pmf {

constexpr int FT = 16;                 // owned tile edge (x and y)
constexpr int FPW = FT + 4;            // LDS prim plane edge (two halo rings)
constexpr int FNP = FPW * FPW;         // 400
constexpr int FQW = FT + 2;            // predicted region edge (one halo ring)
constexpr int FNQ = FQW * FQW;         // 324
constexpr int FNF = (FT + 1) * FT;     // faces per direction in a tile plane: 272
constexpr int FTHREADS = FT * FT + 4 * FT;  // 320

struct FusedGeom {
  int32_t nx, ny, nz, ntx, nty, nchunks, zc;
  int64_t plane, nc;
};

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

template <class T>
__device__ __forceinline__ void load_cons(const T* __restrict__ U, int64_t idx, int64_t nc, T* c) {
#pragma unroll
  for (int v = 0; v < 9; ++v) c[v] = U[v * nc + idx];
}

// Column coordinates (tile-local) of thread t: owned 16x16, then the ring.
__device__ __forceinline__ void column_of(int t, int& ix, int& iy) {
  if (t < FT * FT) {
    ix = t & (FT - 1);
    iy = t >> 4;
    return;
  }
  int h = t - FT * FT;
  if (h < FT) { ix = h; iy = -1; }
  else if (h < 2 * FT) { ix = h - FT; iy = FT; }
  else if (h < 3 * FT) { ix = -1; iy = h - 2 * FT; }
  else { ix = FT; iy = h - 3 * FT; }
}

// Extra cells of the LDS plane: the second halo ring (threads 256..319) and the four
// ring corners (threads 0..3).
__device__ __forceinline__ bool extra_of(int t, int& ix, int& iy) {
  if (t >= FT * FT) {
    int h = t - FT * FT;
    if (h < FT) { ix = h; iy = -2; }
    else if (h < 2 * FT) { ix = h - FT; iy = FT + 1; }
    else if (h < 3 * FT) { ix = -2; iy = h - 2 * FT; }
    else { ix = FT + 1; iy = h - 3 * FT; }
    return true;
  }
  if (t < 4) {
    ix = (t & 1) ? FT : -1;
    iy = (t & 2) ? FT : -1;
    return true;
  }
  return false;
}

__device__ __forceinline__ int pidx(int ix, int iy) { return (iy + 2) * FPW + (ix + 2); }
__device__ __forceinline__ int qidx(int ix, int iy) { return (iy + 1) * FQW + (ix + 1); }

template <class T, int RECON>
__global__ __launch_bounds__(FTHREADS) void k_step_fused(const T* __restrict__ U, T* __restrict__ O, FusedGeom g,
                                                         StepParams<T> P, double* __restrict__ red) {
  __shared__ T sP[9 * FNP];       // primitive plane m (two halo rings)
  __shared__ T sLX[9 * FNQ];      // +x face states, then x-face fluxes
  __shared__ T sLY[9 * FNQ];      // +y face states, then y-face fluxes
  __shared__ double sMax[FTHREADS / 64];

  // XCD-aware tile order: dispatch round-robins workgroups over the 8 XCDs, so
  // logical tile (w % 8) * (nwg / 8) + w / 8 gives each XCD a contiguous tile range.
  const int nwg = gridDim.x;
  int w = blockIdx.x;
  if ((nwg & 7) == 0) w = (w & 7) * (nwg >> 3) + (w >> 3);
  const int tx = w % g.ntx;
  const int ty = (w / g.ntx) % g.nty;
  const int chunk = w / (g.ntx * g.nty);
  const int x0 = tx * FT, y0 = ty * FT;
  const int z0 = chunk * g.zc;
  const int zc = min(g.zc, g.nz - z0);

  const int t = threadIdx.x;
  int ix, iy;
  column_of(t, ix, iy);
  const bool owned = t < FT * FT;
  const int64_t col = (int64_t)wrapi(y0 + iy, g.ny) * g.nx + wrapi(x0 + ix, g.nx);
  int ex = 0, ey = 0;
  const bool has_extra = extra_of(t, ex, ey);
  const int64_t ecol = has_extra ? (int64_t)wrapi(y0 + ey, g.ny) * g.nx + wrapi(x0 + ex, g.nx) : 0;
  const bool xsolver = iy >= 0 && iy < FT && ix >= 0;   // owned + right ring column
  const bool ysolver = ix >= 0 && ix < FT && iy >= 0;   // owned + top ring row

  T Pa[9], Pb[9], Pc[9], Xc[9];
  T Lzp[9], Fzp[9], Dxyp[9];
  {
    T c[9];
    load_cons(U, (int64_t)wrapi(z0 - 2, g.nz) * g.plane + col, g.nc, c);
    cons2prim(c, Pa, P);
    load_cons(U, (int64_t)wrapi(z0 - 1, g.nz) * g.plane + col, g.nc, c);
    cons2prim(c, Pb, P);
  }
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    Lzp[v] = T(0);
    Fzp[v] = T(0);
    Dxyp[v] = T(0);
  }
  double smax = 0.0;

  for (int m = z0 - 1; m <= z0 + zc; ++m) {
    const bool first = m == z0 - 1, last = m == z0 + zc;
    const bool ring_work = !(first || last);  // ring columns feed x/y faces only
    // (1) own column of plane m+1, extra cells of plane m (converted last iteration)
    {
      T c[9];
      load_cons(U, (int64_t)wrapi(m + 1, g.nz) * g.plane + col, g.nc, c);
      cons2prim(c, Pc, P);
    }
    T Ucur[9];
    const bool upd = owned && m - 1 >= z0;
    if (upd) load_cons(U, (int64_t)(m - 1) * g.plane + col, g.nc, Ucur);
#pragma unroll
    for (int v = 0; v < 9; ++v) sP[v * FNP + pidx(ix, iy)] = Pb[v];
    if (has_extra && ring_work) {
#pragma unroll
      for (int v = 0; v < 9; ++v) sP[v * FNP + pidx(ex, ey)] = Xc[v];
    }
    __syncthreads();
    if (has_extra && m + 1 < z0 + zc) {
      T c[9];
      load_cons(U, (int64_t)wrapi(m + 1, g.nz) * g.plane + ecol, g.nc, c);
      cons2prim(c, Xc, P);
    }
    // (2) Hancock prediction of plane m (glm_physics.jl:225-231, integrator_ref.jl:17-26)
    T Rx[9], Ry[9], Rz[9], Lzn[9];
    if (owned || ring_work) {
      T dLx[9], dRx[9], dLy[9], dRy[9], dLz[9], dRz[9];
#pragma unroll
      for (int v = 0; v < 9; ++v) {
        offsets<RECON>(sP[v * FNP + pidx(ix - 1, iy)], Pb[v], sP[v * FNP + pidx(ix + 1, iy)], dLx[v], dRx[v]);
        offsets<RECON>(sP[v * FNP + pidx(ix, iy - 1)], Pb[v], sP[v * FNP + pidx(ix, iy + 1)], dLy[v], dRy[v]);
        offsets<RECON>(Pa[v], Pb[v], Pc[v], dLz[v], dRz[v]);
      }
      T acc[9], fp[9], fm[9], s[9];
#pragma unroll
      for (int v = 0; v < 9; ++v) s[v] = Pb[v] + dRx[v];
      dflux<0>(s, fp, P.igm1);
#pragma unroll
      for (int v = 0; v < 9; ++v) s[v] = Pb[v] + dLx[v];
      dflux<0>(s, fm, P.igm1);
#pragma unroll
      for (int v = 0; v < 9; ++v) acc[v] = fp[v] - fm[v];
#pragma unroll
      for (int v = 0; v < 9; ++v) s[v] = Pb[v] + dRy[v];
      dflux<1>(s, fp, P.igm1);
#pragma unroll
      for (int v = 0; v < 9; ++v) s[v] = Pb[v] + dLy[v];
      dflux<1>(s, fm, P.igm1);
#pragma unroll
      for (int v = 0; v < 9; ++v) acc[v] += fp[v] - fm[v];
#pragma unroll
      for (int v = 0; v < 9; ++v) s[v] = Pb[v] + dRz[v];
      dflux<2>(s, fp, P.igm1);
#pragma unroll
      for (int v = 0; v < 9; ++v) s[v] = Pb[v] + dLz[v];
      dflux<2>(s, fm, P.igm1);
#pragma unroll
      for (int v = 0; v < 9; ++v) acc[v] += fp[v] - fm[v];
      T uh[9], mh[9];
      prim2cons(Pb, uh, P);
#pragma unroll
      for (int v = 0; v < 9; ++v) uh[v] -= P.hdtdx * acc[v];
      cons2prim(uh, mh, P);
      // face states (integrator_ref.jl:27,74-82): +d edge feeds the upper face as its
      // left state, -d edge the lower face as its right state
#pragma unroll
      for (int v = 0; v < 9; ++v) {
        sLX[v * FNQ + qidx(ix, iy)] = mh[v] + dRx[v];
        sLY[v * FNQ + qidx(ix, iy)] = mh[v] + dRy[v];
        Rx[v] = mh[v] + dLx[v];
        Ry[v] = mh[v] + dLy[v];
        Rz[v] = mh[v] + dLz[v];
        Lzn[v] = mh[v] + dRz[v];
      }
    }
    __syncthreads();
    // (3) faces: x and y faces of plane m, z face m-1/2
    const bool xy = !(first || last);
    T Fx[9], Fy[9], Fz[9];
    if (xy && xsolver) {
      T L[9];
#pragma unroll
      for (int v = 0; v < 9; ++v) L[v] = sLX[v * FNQ + qidx(ix - 1, iy)];
      face_flux<0, false>(L, Rx, Fx, P);
    }
    if (xy && ysolver) {
      T L[9];
#pragma unroll
      for (int v = 0; v < 9; ++v) L[v] = sLY[v * FNQ + qidx(ix, iy - 1)];
      face_flux<1, false>(L, Ry, Fy, P);
    }
    if (owned && !first) face_flux<2, false>(Lzp, Rz, Fz, P);
    __syncthreads();
    if (xy && xsolver) {
#pragma unroll
      for (int v = 0; v < 9; ++v) sLX[v * FNF + iy * (FT + 1) + ix] = Fx[v];
    }
    if (xy && ysolver) {
#pragma unroll
      for (int v = 0; v < 9; ++v) sLY[v * FNF + iy * FT + ix] = Fy[v];
    }
    __syncthreads();
    if (owned) {
      // (4) conservative update of plane m-1 (integrator_ref.jl:140-158)
      if (upd) {
        T o[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) o[v] = Ucur[v] + P.dtdx * (Dxyp[v] + (Fzp[v] - Fz[v]));
        o[8] *= P.decay;
        const int64_t idx = (int64_t)(m - 1) * g.plane + col;
#pragma unroll
        for (int v = 0; v < 9; ++v) O[v * g.nc + idx] = o[v];
        smax = nanmax(smax, (double)cell_wavespeed(o, P));
      }
      if (xy) {
#pragma unroll
        for (int v = 0; v < 9; ++v)
          Dxyp[v] = (sLX[v * FNF + iy * (FT + 1) + ix] - sLX[v * FNF + iy * (FT + 1) + ix + 1]) +
                    (sLY[v * FNF + iy * FT + ix] - sLY[v * FNF + (iy + 1) * FT + ix]);
      }
      if (!first) {
#pragma unroll
        for (int v = 0; v < 9; ++v) Fzp[v] = Fz[v];
      }
#pragma unroll
      for (int v = 0; v < 9; ++v) Lzp[v] = Lzn[v];
    }
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      Pa[v] = Pb[v];
      Pb[v] = Pc[v];
    }
  }
  // fold the workgroup's CFL maximum (timestep.jl:8-12 on the new state)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) smax = nanmax(smax, __shfl_xor(smax, o, 64));
  if ((t & 63) == 0) sMax[t >> 6] = smax;
  __syncthreads();
  if (t == 0) {
    double r = sMax[0];
    for (int i = 1; i < FTHREADS / 64; ++i) r = nanmax(r, sMax[i]);
    red[blockIdx.x] = r;
  }
}

__global__ __launch_bounds__(256) void k_fold_max_flat(const double* __restrict__ red, int nb, double* __restrict__ out) {
  __shared__ double sh[4];
  double m = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) m = nanmax(m, red[b]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = nanmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = nanmax(nanmax(sh[0], sh[1]), nanmax(sh[2], sh[3]));
}

static FusedGeom fused_geom(const Grid& G, int zc) {
  FusedGeom g;
  g.nx = G.nx;
  g.ny = G.ny;
  g.nz = G.nz;
  g.ntx = G.nx / FT;
  g.nty = G.ny / FT;
  g.zc = zc;
  g.nchunks = (G.nz + zc - 1) / zc;
  g.plane = (int64_t)G.nx * G.ny;
  g.nc = G.nc;
  return g;
}

bool fused_supported(const Grid& G, int precision, int recon, int rsolve) {
  (void)precision;
  (void)recon;
  (void)rsolve;
  return G.bc[0] == BC_PERIODIC && G.bc[1] == BC_PERIODIC && G.bc[2] == BC_PERIODIC && G.nx % FT == 0 &&
         G.ny % FT == 0 && G.nz >= 2;
}

// z-chunk length: long enough that the two extra prediction planes per chunk stay
// small, short enough to give a few workgroups per CU.
static int pick_zc(const Grid& G) {
  int64_t tiles = (int64_t)(G.nx / FT) * (G.ny / FT);
  int zc = 32;
  while (zc > 4 && tiles * ((G.nz + zc - 1) / zc) < 2048) zc >>= 1;
  return zc < G.nz ? zc : G.nz;
}

int64_t fused_blocks(const Grid& G) {
  if (G.nx % FT || G.ny % FT) return 0;
  FusedGeom g = fused_geom(G, pick_zc(G));
  return (int64_t)g.ntx * g.nty * g.nchunks;
}

template <class T>
bool launch_step_fused(const T* U, T* O, const Grid& G, const StepParams<T>& P, double* red, double* dsmax,
                       hipStream_t st) {
  if (!fused_supported(G, sizeof(T) == 8, P.recon, P.rsolve)) return false;
  FusedGeom g = fused_geom(G, pick_zc(G));
  unsigned nb = (unsigned)((int64_t)g.ntx * g.nty * g.nchunks);
  dim3 blk(FTHREADS), grd(nb);
  if (P.recon == RECON_PPM)
    hipLaunchKernelGGL((k_step_fused<T, RECON_PPM>), grd, blk, 0, st, U, O, g, P, red);
  else if (P.recon == RECON_PLM_MINMOD)
    hipLaunchKernelGGL((k_step_fused<T, RECON_PLM_MINMOD>), grd, blk, 0, st, U, O, g, P, red);
  else
    hipLaunchKernelGGL((k_step_fused<T, RECON_PLM>), grd, blk, 0, st, U, O, g, P, red);
  hipLaunchKernelGGL(k_fold_max_flat, dim3(1), dim3(256), 0, st, red, (int)nb, dsmax);
  return true;
}

template bool launch_step_fused<float>(const float*, float*, const Grid&, const StepParams<float>&, double*, double*,
                                       hipStream_t);
template bool launch_step_fused<double>(const double*, double*, const Grid&, const StepParams<double>&, double*,
                                        double*, hipStream_t);

}  // namespace pmf
