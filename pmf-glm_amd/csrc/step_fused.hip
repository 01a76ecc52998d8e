// Fused, z-marching GLM step for periodic grids (the throughput path; the MI355X
// counterpart of lib/MHDKernels/src/integrator_cube.jl:295-476, same unsplit scheme).
//
// Work decomposition. A workgroup owns a 16x12 column tile in x-y and marches it
// through a chunk of z planes. Its 256 threads (4 wave64s) are bound for the whole
// march to one column each: 192 owned columns plus the 56 columns of the one-cell halo
// ring (corners excluded), which need their Hancock prediction for the tile's boundary
// faces (8 threads have no column). Two LDS planes hold the primitive states of planes
// m and m-1 with both halo rings; the next plane enters the buffer of m-1 as soon as
// each column's own slot there is no longer needed. With the kernel's register budget
// (about 165 VGPRs) a SIMD holds three waves for the float minmod and MonCen-HLL/LLF
// variants (two for the rest), so three (two) 4-wave workgroups share a CU. Per plane m:
//   1. convert the next plane of the own column (and, for a few threads, of the extra
//      halo cells) from conserved to primitive;
//   2. Hancock-predict plane m in every column (cons2prim of the half-step state);
//   3. solve each x and y face of the tile once (L state through LDS, R in registers)
//      and the z face m-1/2 of each owned column (both states in registers);
//   4. update plane m-1 of the owned columns with the reference's flux grouping, damp
//      psi, store, and fold the new cell's CFL speed into a running maximum.
// Every cell is predicted once per tile pass (1.29x with the ring), every face is solved
// once, and the global state is read once and written once per step. Workgroups are
// remapped so that consecutive tiles land on the same XCD (and its L2), which then
// serves the halo re-reads.
#include <cstdlib>
#include <type_traits>

#include "launch.hpp"
#include "surrogate.hpp"

namespace pmf {

constexpr int TX = kFusedTX, TY = kFusedTY;  // owned tile (x, y)
constexpr int NOWN = TX * TY;          // 192
constexpr int NRING = 2 * TX + 2 * TY; // 56 (one-cell ring without corners)
constexpr int FPW = TX + 4;            // LDS prim plane width (two halo rings)
constexpr int FNP = FPW * (TY + 4);    // 320
constexpr int FQW = TX + 2;            // predicted region width (one halo ring)
constexpr int FNQ = FQW * (TY + 2);    // 252
constexpr int FNFX = (TX + 1) * TY;    // x faces of a tile plane: 204
constexpr int FNFY = TX * (TY + 1);    // y faces of a tile plane: 208
constexpr int FTHREADS = 256;
constexpr int FQCAP = NOWN + FNFX + FNFY;  // faces of one tile plane (z, x, y): the surrogate queue

__device__ __forceinline__ int wrapi(int i, int n) {
  int r = i % n;
  return r < 0 ? r + n : r;
}

// Source index of a possibly out-of-range logical index along one axis (boundary.jl:17-31,
// 0-based): periodic wrap, outflow clamp, reflecting mirror about the boundary face with
// flip set (the normal velocity and field change sign). The mirror is clamped into the
// axis as well: only the rows of a ragged last tile row, which are computed but never
// stored, reach past the two ghost layers.
__device__ __forceinline__ int src_index(int i, int n, int bc, bool& flip) {
  flip = false;
  if (bc == BC_PERIODIC) return wrapi(i, n);
  if (i >= 0 && i < n) return i;
  int r;
  if (bc == BC_OUTFLOW) {
    r = i < 0 ? 0 : n - 1;
  } else {
    flip = true;
    r = i < 0 ? -1 - i : 2 * n - 1 - i;
  }
  return min(max(r, 0), n - 1);
}

// Address of cell idx of one field. NARROW: the byte offset fits 32 bits (fields below
// 4 GiB), so every access is the field's scalar base plus one 32-bit vector offset shared
// by the nine fields (global_load/store with saddr) instead of a 64-bit address each.
template <bool NARROW, class T> __device__ __forceinline__ T* cell_at(T* base, int64_t idx) {
  if constexpr (NARROW) {
    using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return reinterpret_cast<T*>(reinterpret_cast<B*>(base) + (uint32_t)idx * (uint32_t)sizeof(T));
  } else {
    return base + idx;
  }
}

template <bool NARROW, class T>
__device__ __forceinline__ void load_cons(const Fields<const T>& U, int64_t idx, T* c) {
#pragma unroll
  for (int v = 0; v < 9; ++v) c[v] = *cell_at<NARROW>(U.v[v], idx);
}

// Ring cell h (0..NRING-1) at distance r from the tile: bottom row, top row, left
// column, right column.
__device__ __forceinline__ void ring_cell(int h, int r, int& ix, int& iy) {
  if (h < TX) { ix = h; iy = -r; }
  else if (h < 2 * TX) { ix = h - TX; iy = TY - 1 + r; }
  else if (h < 2 * TX + TY) { ix = -r; iy = h - 2 * TX; }
  else { ix = TX - 1 + r; iy = h - 2 * TX - TY; }
}

// Column of thread t: owned 16x12, then the ring; threads past the ring have none.
// CORNER (the CT step that forms its edge EMFs): thread NOWN + NRING also owns the
// lower-left corner column (-1, -1), whose cell EMF and +x / +y face states the edges of
// the tile's first row and column use.
template <bool CORNER = false> __device__ __forceinline__ bool column_of(int t, int& ix, int& iy) {
  if (t < NOWN) {
    ix = t % TX;
    iy = t / TX;
    return true;
  }
  if (t < NOWN + NRING) {
    ring_cell(t - NOWN, 1, ix, iy);
    return true;
  }
  if (CORNER && t == NOWN + NRING) {
    ix = -1;
    iy = -1;
    return true;
  }
  ix = 0;
  iy = 0;
  return false;
}

// Extra cells of the LDS plane: the second halo ring (threads NOWN..NOWN+55) and the
// four corners of the first ring (the next four threads). CORNER: the corner (-1, -1) is
// a column, and the two cells its slopes need beyond the rings, (-2, -1) and (-1, -2),
// are extra cells instead.
template <bool CORNER = false> __device__ __forceinline__ bool extra_of(int t, int& ix, int& iy) {
  if (t >= NOWN && t < NOWN + NRING) {
    ring_cell(t - NOWN, 2, ix, iy);
    return true;
  }
  int c = t - NOWN - NRING;
  if (CORNER && c == 0) {
    ix = -2;
    iy = -1;
    return true;
  }
  if (CORNER && c == 4) {
    ix = -1;
    iy = -2;
    return true;
  }
  if (c >= 0 && c < 4) {
    ix = (c & 1) ? TX : -1;
    iy = (c & 2) ? TY : -1;
    return true;
  }
  return false;
}

__device__ __forceinline__ int pidx(int ix, int iy) { return (iy + 2) * FPW + (ix + 2); }
__device__ __forceinline__ int qidx(int ix, int iy) { return (iy + 1) * FQW + (ix + 1); }
// Left-state slots of the x faces (cells ix = -1..TX-1 of the owned rows) and of the y
// faces (cells iy = -1..TY-1 of the owned columns); the slot of a face's left cell is
// also the slot of that face's flux.
__device__ __forceinline__ int xlidx(int ix, int iy) { return iy * (TX + 1) + (ix + 1); }
__device__ __forceinline__ int ylidx(int ix, int iy) { return (iy + 1) * TX + ix; }

// LDS of one workgroup of the fused step (37.9 KB without sources: four workgroups fit
// a CU's 160 KB).
// Where the x/y divergence of a column is carried to the next plane: in LDS for PPM,
// whose edge offsets fill the three-wave register budget (it freed 8 VGPRs and the 16-20
// spilled dwords: 256^3 PPM-HLLD 0.99 -> 0.95 ms, PPM-HLL 0.88 -> 0.82 ms), in registers
// for PLM, where the extra LDS traffic cost 1-2%, and for the surrogate step, whose LDS
// is full.
template <int RECON, bool SURR>
constexpr bool dxy_in_lds() {
  return RECON == RECON_PPM && !SURR;
}

// EDGES (the CT step that forms its edge EMFs): the x faces of ring row -1 and the y
// faces of ring column -1 get slots too, and the z-face fluxes and cell EMFs of the
// predicted region are exchanged through LDS.
constexpr int FNFXE = (TX + 1) * (TY + 1);  // x faces of rows -1..TY-1: 221
constexpr int FNFYE = (TX + 1) * (TY + 1);  // y faces of columns -1..TX-1: 221
template <class T, bool SRC = false, bool DXL = true, bool EDGES = false>
struct FusedLds {
  T sP[2][9 * FNP];  // primitive planes m (current) and m-1, alternating
  T sLX[9 * (EDGES ? FNFXE : FNFX)];  // left states of the x faces, then the x-face fluxes
  T sLY[9 * (EDGES ? FNFYE : FNFY)];  // left states of the y faces, then the y-face fluxes
  T sD[DXL ? 9 * NOWN : 1];  // DXL: each owned column's x/y divergence, carried a plane
  unsigned char sMk[SRC ? FNQ : 1];  // masked-fallback flags of the predicted cells
  T sZf[EDGES ? 3 * FNQ : 1];        // EDGES: mass, Bx and By fluxes of the z face m-1/2
  T sEz[EDGES ? FNQ : 1];            // EDGES: cell EMF z of plane m
  T sExy[EDGES ? 2 * 2 * FNQ : 1];   // EDGES: cell EMFs x, y of planes m and m-1 (by parity)
};

// One work item: tile w (x fastest, then y, then z chunk) of the grid g, marched through
// its z chunk. Called by all FTHREADS threads of the workgroup; returns the calling
// thread's running CFL maximum of the cells it stored. SURR: the surrogate-assisted
// variant (float only): each face is gated by the jump detector and solved by the
// learned flux (wave-level MFMA) or by HLLD; sW holds the network's LDS image
// (surrogate.hpp), sB its biases (global memory), and smooth_count counts the faces that
// used it (each face once, as the lower face of its cell).
// Cell EMF -v x B (integrator_ref.jl:133-138) from the predicted velocity v and the
// start-of-step field b (primitive slots), without contraction, so that the step's two
// forms of the CT update (workspace slots, or edges formed in the step) use equal bits.
template <class T> __device__ __forceinline__ void cell_emf(const T* v, const T* b, T& ex, T& ey, T& ez) {
#pragma clang fp contract(off)
  ex = v[3] * b[6] - v[2] * b[7];
  ey = v[1] * b[7] - v[3] * b[5];
  ez = v[2] * b[5] - v[1] * b[6];
}

// EDGES (CT only): the step also forms the edge EMFs of its stored cells (ct.jl:71-93, the
// GS05 upwind average) from the face fluxes and cell EMFs it has just computed, and writes
// them to workspace slots 15-17 instead of writing slots 3-14 for a separate edge pass.
template <class T, int RECON, int RS, bool SURR, bool SRC = false, bool NARROW = false, bool CT = false,
          bool BCG = false, bool EDGES = false>
__device__ __forceinline__ T fused_item(const Fields<const T>& U, const Fields<T>& O, const FusedGeom& g,
                                        const StepParams<T>& P, int w,
                                        FusedLds<T, SRC, dxy_in_lds<RECON, SURR>(), EDGES>& S,
                                        const unsigned short* __restrict__ sW, const float* __restrict__ sB,
                                        float thr, unsigned& smooth_count, const SrcParams<T>& src,
                                        SurrQueue<FQCAP>* Qp = nullptr, T* __restrict__ ctp = nullptr) {
  static_assert(!(CT && SURR), "the fused CT step has no surrogate");
  static_assert(!(BCG && (CT || SURR)), "the CT and surrogate steps run on periodic grids");
  static_assert(!EDGES || (CT && !SRC), "edge EMFs belong to the plain CT step");
  const int tx = w % g.ntx;
  const int ty = (w / g.ntx) % g.nty;
  const int chunk = w / (g.ntx * g.nty);
  const int x0 = tx * TX, y0 = ty * TY;
  const int z0 = chunk * g.zc;
  const int zc = min(g.zc, g.nz - z0);

  const int t = threadIdx.x;
  int ix, iy;
  const bool has_col = column_of<EDGES>(t, ix, iy);
  const bool owned = t < NOWN;
  // ragged last tile column / row: computed (their cells feed no stored cell), not stored
  const bool stores = owned && x0 + ix < g.nx && y0 + iy < g.ny;
  // live extent of the tile: a ragged last tile (fewer than TX columns or TY rows left in
  // the grid) solves only the faces of its stored cells and predicts only their states, so
  // the waves over its dead rows skip their work
  const int vx = min(TX, g.nx - x0), vy = min(TY, g.ny - y0);
  // the column's periodic image: its source cells on periodic grids, and the index of the
  // radiation correction buffer on every grid (integrator_ref.jl:50-53)
  const int64_t pcol = (int64_t)wrapi(y0 + iy, g.ny) * g.nx + wrapi(x0 + ix, g.nx);
  int ex = 0, ey = 0;
  const bool has_extra = extra_of<EDGES>(t, ex, ey);
  int64_t col, ecol = 0;
  int flips = 0;  // BCG: bit 0/1 = x/y mirror of the column, bit 2/3 = of the extra cell
  if constexpr (BCG) {
    // source cells under outflow / reflecting boundaries (boundary.jl:17-31)
    bool fa, fb;
    col = (int64_t)src_index(y0 + iy, g.ny, g.bc[1], fb) * g.nx + src_index(x0 + ix, g.nx, g.bc[0], fa);
    flips = (fa ? 1 : 0) | (fb ? 2 : 0);
    if (has_extra) {
      ecol = (int64_t)src_index(y0 + ey, g.ny, g.bc[1], fb) * g.nx + src_index(x0 + ex, g.nx, g.bc[0], fa);
      flips |= (fa ? 4 : 0) | (fb ? 8 : 0);
    }
  } else {
    col = pcol;
    if (has_extra) ecol = (int64_t)wrapi(y0 + ey, g.ny) * g.nx + wrapi(x0 + ex, g.nx);
  }
  // conserved state of logical plane z of a column (sh = 0) or of the extra cell (sh = 2);
  // BCG: the source plane and the sign flips of the reflected normal components
  // (boundary.jl:34-41)
  auto load_plane = [&](int z, int64_t c2, int sh, T* c) {
    if constexpr (BCG) {
      bool fz = false;
      int kz = z;
      if (z < 0 || z >= g.nz) kz = src_index(z, g.nz, g.bc[2], fz);
      load_cons<NARROW>(U, (int64_t)kz * g.plane + c2, c);
      const int f = (flips >> sh) & 3;
      if (f & 1) { c[1] = -c[1]; c[5] = -c[5]; }
      if (f & 2) { c[2] = -c[2]; c[6] = -c[6]; }
      if (fz) { c[3] = -c[3]; c[7] = -c[7]; }
    } else {
      (void)sh;
      load_cons<NARROW>(U, (int64_t)wrapi(z, g.nz) * g.plane + c2, c);
    }
  };
  const bool xsolver = has_col && iy >= 0 && iy < vy && ix >= 0 && ix <= vx;  // owned + right ring column
  const bool ysolver = has_col && ix >= 0 && ix < vx && iy >= 0 && iy <= vy;  // owned + top ring row
  // columns whose predicted states feed a live face (every column of a whole tile)
  const bool pred_live =
      has_col && ((ix >= -1 && ix <= vx && iy >= 0 && iy < vy) || (ix >= 0 && ix < vx && iy >= -1 && iy <= vy));
  const bool xleft = has_col && ix < TX && iy >= 0 && iy < TY;     // owned + left ring column
  const bool yleft = has_col && iy < TY && ix >= 0 && ix < TX;     // owned + bottom ring row
  // EDGES: x/y face slots for ring row -1 and ring column -1 as well
  constexpr int NFX = EDGES ? FNFXE : FNFX, NFY = EDGES ? FNFYE : FNFY;
  auto xl = [](int a, int b) { return EDGES ? (b + 1) * (TX + 1) + (a + 1) : xlidx(a, b); };
  auto yl = [](int a, int b) { return EDGES ? (b + 1) * (TX + 1) + (a + 1) : ylidx(a, b); };
  // EDGES roles: the live part of ring row -1 and of ring column -1 also solve their z faces
  // and their x (row) or y (column) faces and publish their cell EMFs, which the edges of
  // the tile's first row and column read; the corner column (-1, -1) publishes its cell
  // EMF and the left states of the two faces beside it.
  const bool ebot = EDGES && has_col && iy == -1 && ix >= 0 && ix < vx;
  const bool eleft = EDGES && has_col && ix == -1 && iy >= 0 && iy < vy;
  const bool ecorner = EDGES && has_col && ix == -1 && iy == -1;
  const bool xleft_e = xleft || (EDGES && has_col && iy == -1 && ix < TX);
  const bool yleft_e = yleft || (EDGES && has_col && ix == -1 && iy < TY);
  T fybz_p = T(0), mfy_p = T(0), fxbz_p = T(0), mfx_p = T(0);  // EDGES: own x/y-face fluxes, plane m-1
  const int me = pidx(ix, iy);
  // CT workspace slot q of cell idx (NARROW: the packed buffer's byte offsets fit 32 bits)
  auto ctslot = [&](int q, int64_t idx) -> T& { return *cell_at<NARROW>(ctp + (int64_t)q * g.nc, idx); };

  // Column registers: Lzp = left state of the z face below the current plane, Fzp = flux
  // of the z face below the plane being updated, Dxy = that plane's x/y divergence (PPM:
  // in the thread's own LDS slots, nine registers fewer across the prediction and the z
  // face).
  // The LDS buffer of plane m-1 is refilled with plane m+1 while plane m is processed:
  // each thread overwrites its own column slot (read only by itself, for the z slope)
  // after its prediction, and the extra cells' slots are never read during plane m.
  constexpr bool DXL = dxy_in_lds<RECON, SURR>();
  T Lzp[9], Fzp[9], Dxy[DXL ? 1 : 9];
  auto dxy = [&](int v) -> T& {
    if constexpr (DXL) return S.sD[v * NOWN + t];
    else return Dxy[v];
  };
  {
    T c[9], q[9];
    if (has_col) {
      load_plane(z0 - 2, col, 0, c);
      cons2prim(c, q, P);
#pragma unroll
      for (int v = 0; v < 9; ++v) S.sP[(z0 - 2) & 1][v * FNP + me] = q[v];
      load_plane(z0 - 1, col, 0, c);
      cons2prim(c, q, P);
#pragma unroll
      for (int v = 0; v < 9; ++v) S.sP[(z0 - 1) & 1][v * FNP + me] = q[v];
    }
    // EDGES: ring row -1 and column -1 are predicted on the chunk's first plane as well,
    // so that plane's extra cells (their slopes' second neighbours) are needed too
    if (EDGES && has_extra) {
      load_plane(z0 - 1, ecol, 2, c);
      cons2prim(c, q, P);
#pragma unroll
      for (int v = 0; v < 9; ++v) S.sP[(z0 - 1) & 1][v * FNP + pidx(ex, ey)] = q[v];
    }
  }
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    Lzp[v] = T(0);
    Fzp[v] = T(0);
    if (owned) dxy(v) = T(0);
  }
  T smax = T(0);

  // CT workspace stores are issued after the loads that follow them in program order
  // (a wave's vector-memory counter retires loads and stores in issue order, so a store
  // ahead of a load makes the wait for the load also wait for the store's write): the
  // cell EMF and z-face fluxes of plane m go out after the update's loads of plane m-1,
  // and the x/y-face fluxes of plane m after the next plane's loads.
  T ctc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};  // plane m: ecx ecy ecz, mfz fzbx fzby
  bool ct_emf = false, ct_zf = false;
  T ctxy[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};  // the previous plane: mfx fxby fxbz, mfy fybx fybz
  auto ct_flush_cell = [&](int m) {
    if constexpr (CT) {
      const int64_t idx = (int64_t)wrapi(m, g.nz) * g.plane + col;
      if (ct_emf) {
        ctslot(12, idx) = ctc[0];
        ctslot(13, idx) = ctc[1];
        ctslot(14, idx) = ctc[2];
      }
      if (ct_zf) {
        ctslot(5, idx) = ctc[3];
        ctslot(10, idx) = ctc[4];
        ctslot(11, idx) = ctc[5];
      }
      ct_emf = ct_zf = false;
    }
  };

  // (4) the update of plane m-1 once the z face m-1/2 is known (integrator_ref.jl:140-158);
  // shifts the z-face carry. Owned columns only.
  auto zface_update = [&](const T* Fz, const T* Lzn, int m, bool first) {
    bool flushed = false;
    if (m - 1 >= z0 && stores) {
      T o[9];
      const int64_t idx = (int64_t)(m - 1) * g.plane + col;
      load_cons<NARROW>(U, idx, o);
      if constexpr (CT) {
        ct_flush_cell(m);
        flushed = true;
      }
      if constexpr (SRC) {
        // tracked increments, drag propagator, admissibility mark (integrator_cube.jl:423-441)
        T du[9], U0[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) {
          U0[v] = o[v];
          du[v] = P.dtdx * (dxy(v) + (Fzp[v] - Fz[v]));
        }
        src_update(U0, du, o, src, idx, P, CT);
      } else {
#pragma unroll
        for (int v = 0; v < 9; ++v) o[v] += P.dtdx * (dxy(v) + (Fzp[v] - Fz[v]));
      }
      if constexpr (CT) {
        o[8] = T(0);  // no cleaning field under CT (integrator_ref.jl:150-158)
      } else {
        o[8] *= P.decay;
      }
#pragma unroll
      for (int v = 0; v < 9; ++v) *cell_at<NARROW>(O.v[v], idx) = o[v];
      // CT: the cell field is replaced by the face average after the step, so the CFL
      // speed is taken from the final state instead
      if constexpr (!CT) smax = nanmax(smax, cell_wavespeed(o, P));
    }
    if constexpr (CT) {
      if (!flushed) ct_flush_cell(m);
    }
    if (!first) {
#pragma unroll
      for (int v = 0; v < 9; ++v) Fzp[v] = Fz[v];
    }
#pragma unroll
    for (int v = 0; v < 9; ++v) Lzp[v] = Lzn[v];
  };

  static_assert(!(SRC && SURR), "the surrogate path has no source terms");
  bool mkz = false, mkzp = false;  // masked-fallback flags of this column's planes m and m-1
  // PLM without source terms: this column's z difference q(m) - q(m-1), carried from
  // plane to plane (the source and surrogate variants, short of registers, re-read it)
  constexpr bool ZCARRY = RECON != RECON_PPM && !SRC && !SURR;
  T dzc[9];
  for (int m = z0 - 1; m <= z0 + zc; ++m) {
    const bool first = m == z0 - 1, last = m == z0 + zc;
    const bool ring_work = !(first || last);  // ring columns feed x/y faces only
    // the x and y faces this thread solves on plane m (EDGES: on the first plane the stored
    // columns' own faces too, whose fluxes the edges of the next plane read)
    const bool xs = EDGES ? (ring_work ? (xsolver || ebot) : (first && stores)) : (ring_work && xsolver);
    const bool ys = EDGES ? (ring_work ? (ysolver || eleft) : (first && stores)) : (ring_work && ysolver);
    const bool ezf = EDGES && (ebot || eleft) && !last;  // EDGES: a ring column's z face and carry
    const T* __restrict__ cur = S.sP[m & 1];
    T* __restrict__ nxt = S.sP[(m + 1) & 1];   // holds plane m-1 now, plane m+1 after this pass
    // (1) plane m+1 of the own column and of the extra cell (in flight during the barrier)
    T Pn[9], Xc[9];
    if (has_col) load_plane(m + 1, col, 0, Pn);
    const bool extra_now = has_extra && m + 1 < z0 + zc;
    if (extra_now) load_plane(m + 1, ecol, 2, Xc);
    // SRC: the radiation correction and the fallback mask of this plane's predicted cell,
    // also in flight during the barrier (they are used after the prediction)
    T cv[3] = {T(0), T(0), T(0)};
    T mkv = T(0);
    if constexpr (SRC) {
      if (has_col) {
        const int64_t pc = (int64_t)wrapi(m, g.nz) * g.plane + pcol;
        if (src.flags & SRC_RAD) {
#pragma unroll
          for (int c = 0; c < 3; ++c) cv[c] = src.corr[(1 + c) * src.stride + pc];
        }
        if (src.check == CHECK_MASKED) mkv = src.corr[4 * src.stride + pc];
      }
    }
    // CT: the face fields this thread's solves take as the normal field (the z face m-1/2
    // of an owned column, the x and y faces of plane m), also in flight during the barrier
    T bfz = T(0), bfx = T(0), bfy = T(0);
    if constexpr (CT) {
      if ((stores || ezf) && !first) bfz = ctslot(2, (int64_t)wrapi(m, g.nz) * g.plane + col);
      if (xs) bfx = ctslot(0, (int64_t)wrapi(m, g.nz) * g.plane + col);
      if (ys) bfy = ctslot(1, (int64_t)wrapi(m, g.nz) * g.plane + col);
      // the x/y-face fluxes of plane m-1 of the chunk's own cells, now behind the loads
      if (!EDGES && stores && m - 1 >= z0 && m - 1 < z0 + zc) {
        const int64_t idx = (int64_t)(m - 1) * g.plane + col;
        ctslot(3, idx) = ctxy[0];
        ctslot(6, idx) = ctxy[1];
        ctslot(7, idx) = ctxy[2];
        ctslot(4, idx) = ctxy[3];
        ctslot(8, idx) = ctxy[4];
        ctslot(9, idx) = ctxy[5];
      }
    }
    if constexpr (SURR) {
      if (t == 0) Qp->n = 0u;  // the previous plane's queue was last read before its final barrier
    }
    __syncthreads();
    if (extra_now) {
      T q[9];
      cons2prim(Xc, q, P);
#pragma unroll
      for (int v = 0; v < 9; ++v) nxt[v * FNP + pidx(ex, ey)] = q[v];
    }
    if (has_col) {
      T q[9];
      cons2prim(Pn, q, P);
#pragma unroll
      for (int v = 0; v < 9; ++v) Pn[v] = q[v];
    }
    // (2) Hancock prediction of plane m (glm_physics.jl:225-231, integrator_ref.jl:17-26)
    T Rx[9], Ry[9], Rz[9], Lzn[9];
    if (stores || (pred_live && ring_work) || ezf || (ecorner && ring_work)) {
      T m0[9], dLx[9], dRx[9], dLy[9], dRy[9], dLz[9], dRz[9];
      if constexpr (ZCARRY) {
        if (first) {
#pragma unroll
          for (int v = 0; v < 9; ++v) dzc[v] = cur[v * FNP + me] - nxt[v * FNP + me];
        }
      }
#pragma unroll
      for (int v = 0; v < 9; ++v) {
        m0[v] = cur[v * FNP + me];
        offsets<RECON>(cur[v * FNP + me - 1], m0[v], cur[v * FNP + me + 1], dLx[v], dRx[v]);
        offsets<RECON>(cur[v * FNP + me - FPW], m0[v], cur[v * FNP + me + FPW], dLy[v], dRy[v]);
        if constexpr (ZCARRY) {
          // the lower z difference of plane m is the upper one of plane m-1
          const T dr = Pn[v] - m0[v];
          offsets_diff<RECON>(dzc[v], dr, dLz[v], dRz[v]);
          dzc[v] = dr;
        } else {
          offsets<RECON>(nxt[v * FNP + me], m0[v], Pn[v], dLz[v], dRz[v]);
        }
        nxt[v * FNP + me] = Pn[v];
      }
      if constexpr (SRC) {
        // masked / global fallback: flagged cells are predicted with zero slopes
        // (integrator_cube.jl:330-335); mkv is the predicted cell's slot 4
        const bool mk = src.check == CHECK_GLOBAL || (src.check == CHECK_MASKED && mkv > T(0));
        if (mk) {
#pragma unroll
          for (int v = 0; v < 9; ++v) dLx[v] = dRx[v] = dLy[v] = dRy[v] = dLz[v] = dRz[v] = T(0);
        }
        S.sMk[qidx(ix, iy)] = mk;
        mkz = mk;
      }
      T uh[9], mh[9];
      if constexpr (RECON != RECON_PPM) {
        // PLM edges are m0 -/+ d: half of each flux difference straight from its even and
        // odd parts (shared with the per-cell step)
        hancock_plm(m0, dRx, dRy, dRz, uh, P);
      } else {
        T acc[9], fp[9], fm[9], s[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) s[v] = m0[v] + dRx[v];
        dflux<0>(s, fp, P.igm1);
#pragma unroll
        for (int v = 0; v < 9; ++v) s[v] = m0[v] + dLx[v];
        dflux<0>(s, fm, P.igm1);
#pragma unroll
        for (int v = 0; v < 9; ++v) acc[v] = fp[v] - fm[v];
#pragma unroll
        for (int v = 0; v < 9; ++v) s[v] = m0[v] + dRy[v];
        dflux<1>(s, fp, P.igm1);
#pragma unroll
        for (int v = 0; v < 9; ++v) s[v] = m0[v] + dLy[v];
        dflux<1>(s, fm, P.igm1);
#pragma unroll
        for (int v = 0; v < 9; ++v) acc[v] += fp[v] - fm[v];
#pragma unroll
        for (int v = 0; v < 9; ++v) s[v] = m0[v] + dRz[v];
        dflux<2>(s, fp, P.igm1);
#pragma unroll
        for (int v = 0; v < 9; ++v) s[v] = m0[v] + dLz[v];
        dflux<2>(s, fm, P.igm1);
#pragma unroll
        for (int v = 0; v < 9; ++v) acc[v] += fp[v] - fm[v];
        prim2cons(m0, uh, P);
#pragma unroll
        for (int v = 0; v < 9; ++v) uh[v] -= P.hdtdx * acc[v];
      }
      cons2prim(uh, mh, P);
      // face states (integrator_ref.jl:27,74-82): the +d edge is the left state of the
      // upper face, the -d edge the right state of the lower face
      if constexpr (SRC) {
        // face states with the drag propagation and the radiation correction
        // (integrator_cube.jl:342-374, integrator_ref.jl:41-48)
        T fx[9], fy[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) {
          fx[v] = mh[v] + dRx[v];
          fy[v] = mh[v] + dRy[v];
          Rx[v] = mh[v] + dLx[v];
          Ry[v] = mh[v] + dLy[v];
          Rz[v] = mh[v] + dLz[v];
          Lzn[v] = mh[v] + dRz[v];
        }
        src_face_pre(fx, m0, dRx, src, cv);
        src_face_pre(fy, m0, dRy, src, cv);
        src_face_pre(Rx, m0, dLx, src, cv);
        src_face_pre(Ry, m0, dLy, src, cv);
        src_face_pre(Rz, m0, dLz, src, cv);
        src_face_pre(Lzn, m0, dRz, src, cv);
        if (xleft) {
#pragma unroll
          for (int v = 0; v < 9; ++v) S.sLX[v * NFX + xl(ix, iy)] = fx[v];
        }
        if (yleft) {
#pragma unroll
          for (int v = 0; v < 9; ++v) S.sLY[v * NFY + yl(ix, iy)] = fy[v];
        }
      } else {
        if (xleft_e) {
#pragma unroll
          for (int v = 0; v < 9; ++v) S.sLX[v * NFX + xl(ix, iy)] = mh[v] + dRx[v];
        }
        if (yleft_e) {
#pragma unroll
          for (int v = 0; v < 9; ++v) S.sLY[v * NFY + yl(ix, iy)] = mh[v] + dRy[v];
        }
#pragma unroll
        for (int v = 0; v < 9; ++v) {
          Rx[v] = mh[v] + dLx[v];
          Ry[v] = mh[v] + dLy[v];
          Rz[v] = mh[v] + dLz[v];
        }
#pragma unroll
        for (int v = 0; v < 9; ++v) Lzn[v] = mh[v] + dRz[v];
      }
      if constexpr (CT) {
        // cell EMF -v x B from the predicted velocity (SRC: with the drag propagation and
        // radiation correction of a zero-offset face state) and the start-of-step field
        // (integrator_ref.jl:133-138), for the chunk's own cells
        if constexpr (EDGES) {
          // cell EMFs of the stored cells, of ring row -1 and column -1 (also on the first
          // plane: the edges of plane z0 read plane z0-1's) and of the corner, by plane parity
          if (((stores || ebot || eleft) && !last) || (ecorner && ring_work)) {
            const int q = qidx(ix, iy);
            T* E = S.sExy + (m & 1) * 2 * FNQ;
            cell_emf(mh, m0, E[q], E[FNQ + q], S.sEz[q]);
          }
        } else if (stores && !first && !last) {
          T qc[9];
#pragma unroll
          for (int v = 0; v < 9; ++v) qc[v] = mh[v];
          if constexpr (SRC) {
            const T zero[9] = {};
            src_face_pre(qc, m0, zero, src, cv);
          }
          cell_emf(qc, m0, ctc[0], ctc[1], ctc[2]);
          ct_emf = true;
        }
      }
      // (3a) z face m-1/2 (both states in registers), then (4) the update of plane m-1,
      // whose inputs all exist now (integrator_ref.jl:140-158)
      if constexpr (!SURR) {
        if (stores) {
          T Fz[9];
          if (!first) {
            if constexpr (CT) {
              // riemann_ct on the z face m-1/2 = the lower face of plane m: the face field
              // in both states; its mass and transverse-B fluxes go to the workspace
              // (integrator_ref.jl:83-132) when plane m belongs to this chunk
              Lzp[7] = bfz;
              Rz[7] = bfz;
              face_flux<2, true, T, RS>(Lzp, Rz, Fz, P);
              if constexpr (EDGES) {
                if (!last) {
                  const int q = qidx(ix, iy);
                  S.sZf[q] = Fz[0];
                  S.sZf[FNQ + q] = Fz[5];
                  S.sZf[2 * FNQ + q] = Fz[6];
                }
              } else if (stores && !last) {
                ctc[3] = Fz[0];
                ctc[4] = Fz[5];
                ctc[5] = Fz[6];
                ct_zf = true;
              }
            } else if (SRC && src.check == CHECK_MASKED && (mkz || mkzp)) {
              face_flux<2, false, T, RS_LLF>(Lzp, Rz, Fz, P);
            } else {
              face_flux<2, false, T, RS>(Lzp, Rz, Fz, P);
            }
          }
          zface_update(Fz, Lzn, m, first);
        } else if (ezf) {
          if constexpr (EDGES) {
            // ring row -1 / column -1: the z face m-1/2 for the edges, and the carry
            if (!first) {
              T Fz[9];
              Lzp[7] = bfz;
              Rz[7] = bfz;
              face_flux<2, true, T, RS>(Lzp, Rz, Fz, P);
              const int q = qidx(ix, iy);
              S.sZf[q] = Fz[0];
              S.sZf[FNQ + q] = Fz[5];
              S.sZf[2 * FNQ + q] = Fz[6];
            }
#pragma unroll
            for (int v = 0; v < 9; ++v) Lzp[v] = Lzn[v];
          }
        }
      }
      mkzp = mkz;
    }
    else if (has_col) {
      // ring column with no prediction on this plane: still deposit plane m+1 (and carry
      // the column's z difference)
#pragma unroll
      for (int v = 0; v < 9; ++v) {
        if constexpr (ZCARRY) dzc[v] = Pn[v] - cur[v * FNP + me];
        nxt[v * FNP + me] = Pn[v];
      }
    }
    __syncthreads();
    // (3b) x and y faces of plane m, each solved once
    const bool xy = EDGES ? !last : !(first || last);
    if constexpr (SURR) {
      // Surrogate-assisted faces of plane m: (a) every face of the plane (z, x, y) is
      // gated; rough faces take HLLD now and smooth ones queue their network inputs;
      // (b) the waves evaluate the queue in dense batches of 32 (one MFMA pass each);
      // (c) each smooth face adds its learned dissipation to its central flux. Face
      // states were read in (a), so the fluxes may overwrite them after (b)'s barriers.
      SurrQueue<FQCAP>& Q = *Qp;
      const bool zf = stores && !first, xf = xsolver && xy, yf = ysolver && xy;
      float fz[9], sz[9], fx[9], sx[9], fy[9], sy[9];
      int qz = -1, qx = -1, qy = -1;
      if (zf) {
        float xin[SURR_IN];
        const bool sm = surr_face_begin<2>(Lzp, Rz, P, thr, true, fz, sz, xin);
        qz = surr_enqueue(sm, xin, Q);
      }
      if (xf) {
        float L[9], xin[SURR_IN];
#pragma unroll
        for (int v = 0; v < 9; ++v) L[v] = S.sLX[v * FNFX + xlidx(ix - 1, iy)];
        const bool sm = surr_face_begin<0>(L, Rx, P, thr, true, fx, sx, xin);
        qx = surr_enqueue(sm, xin, Q);
      }
      if (yf) {
        float L[9], xin[SURR_IN];
#pragma unroll
        for (int v = 0; v < 9; ++v) L[v] = S.sLY[v * FNFY + ylidx(ix, iy - 1)];
        const bool sm = surr_face_begin<1>(L, Ry, P, thr, true, fy, sy, xin);
        qy = surr_enqueue(sm, xin, Q);
      }
      __syncthreads();
      const int nq = (int)Q.n;
      for (int b = (t >> 6) * 32; b < nq; b += FTHREADS / 2) surr_batch(b, nq, Q, sW, sB);
      __syncthreads();
      float y[SURR_OUT];
      if (stores) {
        T Fz[9];
        if (zf) {
          surr_result(Q, qz, y);
          surr_face_end<2>(fz, sz, y, qz >= 0, Fz);
          // counts each face once: the lower faces of the stored cells
          smooth_count += qz >= 0 && stores && !last;
        }
        zface_update(Fz, Lzn, m, first);
      }
      if (xy) {
        if (xf) {
          T Fx[9];
          surr_result(Q, qx, y);
          surr_face_end<0>(fx, sx, y, qx >= 0, Fx);
          smooth_count += qx >= 0 && stores;
#pragma unroll
          for (int v = 0; v < 9; ++v) S.sLX[v * FNFX + xlidx(ix - 1, iy)] = Fx[v];
        }
        if (yf) {
          T Fy[9];
          surr_result(Q, qy, y);
          surr_face_end<1>(fy, sy, y, qy >= 0, Fy);
          smooth_count += qy >= 0 && stores;
#pragma unroll
          for (int v = 0; v < 9; ++v) S.sLY[v * FNFY + ylidx(ix, iy - 1)] = Fy[v];
        }
        __syncthreads();
        if (stores) {
#pragma unroll
          for (int v = 0; v < 9; ++v)
            dxy(v) = (S.sLX[v * FNFX + xlidx(ix - 1, iy)] - S.sLX[v * FNFX + xlidx(ix, iy)]) +
                     (S.sLY[v * FNFY + ylidx(ix, iy - 1)] - S.sLY[v * FNFY + ylidx(ix, iy)]);
        }
      }
    } else if (xy) {
      // Each face's left-state slot is read and then overwritten with its flux by the one
      // thread that solves the face, so the fluxes go back to LDS without a barrier.
      {
      T Fx[9], Fy[9];
      // masked fallback: LLF on every face touching a flagged cell (integrator_cube.jl:68-80)
      if (xs) {
        T L[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) L[v] = S.sLX[v * NFX + xl(ix - 1, iy)];
        if constexpr (CT) {
          L[5] = bfx;
          Rx[5] = bfx;
          face_flux<0, true, T, RS>(L, Rx, Fx, P);
          if constexpr (!EDGES) {
            ctxy[0] = Fx[0];
            ctxy[1] = Fx[6];
            ctxy[2] = Fx[7];
          }
        } else if (SRC && src.check == CHECK_MASKED && (S.sMk[qidx(ix - 1, iy)] | S.sMk[qidx(ix, iy)])) {
          face_flux<0, false, T, RS_LLF>(L, Rx, Fx, P);
        } else {
          face_flux<0, false, T, RS>(L, Rx, Fx, P);
        }
#pragma unroll
        for (int v = 0; v < 9; ++v) S.sLX[v * NFX + xl(ix - 1, iy)] = Fx[v];
      }
      if (ys) {
        T L[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) L[v] = S.sLY[v * NFY + yl(ix, iy - 1)];
        if constexpr (CT) {
          L[6] = bfy;
          Ry[6] = bfy;
          face_flux<1, true, T, RS>(L, Ry, Fy, P);
          if constexpr (!EDGES) {
            ctxy[3] = Fy[0];
            ctxy[4] = Fy[5];
            ctxy[5] = Fy[7];
          }
        } else if (SRC && src.check == CHECK_MASKED && (S.sMk[qidx(ix, iy - 1)] | S.sMk[qidx(ix, iy)])) {
          face_flux<1, false, T, RS_LLF>(L, Ry, Fy, P);
        } else {
          face_flux<1, false, T, RS>(L, Ry, Fy, P);
        }
#pragma unroll
        for (int v = 0; v < 9; ++v) S.sLY[v * NFY + yl(ix, iy - 1)] = Fy[v];
      }
      }
      __syncthreads();
      if (stores) {
#pragma unroll
        for (int v = 0; v < 9; ++v)
          dxy(v) = (S.sLX[v * NFX + xl(ix - 1, iy)] - S.sLX[v * NFX + xl(ix, iy)]) +
                   (S.sLY[v * NFY + yl(ix, iy - 1)] - S.sLY[v * NFY + yl(ix, iy)]);
      }
      if constexpr (EDGES) {
        // edge EMFs of the stored cells of plane m (ct.jl:71-93, k_ct_emf's argument lists):
        // this plane's x/y-face fluxes from LDS (each face's flux sits in its left cell's
        // slot), its z-face fluxes and cell EMFs from LDS, plane m-1's own x/y-face fluxes
        // from registers and its cell EMFs from the other parity
        if (stores) {
          const int xo = xl(ix - 1, iy), xjm = xl(ix - 1, iy - 1);
          const int yo = yl(ix, iy - 1), yim = yl(ix - 1, iy - 1);
          const T mfx_c = S.sLX[xo], fxbz_c = S.sLX[7 * NFX + xo];
          const T mfy_c = S.sLY[yo], fybz_c = S.sLY[7 * NFY + yo];
          if (!first) {
            const int q = qidx(ix, iy), qim = qidx(ix - 1, iy), qjm = qidx(ix, iy - 1), qimjm = qidx(ix - 1, iy - 1);
            const T* Ec = S.sExy + (m & 1) * 2 * FNQ;
            const T* Ep = S.sExy + ((m - 1) & 1) * 2 * FNQ;
            const T* Zf = S.sZf;
            const T ez = gs05(-S.sLX[6 * NFX + xo], -S.sLX[6 * NFX + xjm], S.sLY[5 * NFY + yo], S.sLY[5 * NFY + yim],
                              mfx_c, S.sLX[xjm], mfy_c, S.sLY[yim], S.sEz[q], S.sEz[qim], S.sEz[qjm], S.sEz[qimjm]);
            const T exv = gs05(-fybz_c, -fybz_p, Zf[2 * FNQ + q], Zf[2 * FNQ + qjm], mfy_c, mfy_p, Zf[q], Zf[qjm],
                               Ec[q], Ec[qjm], Ep[q], Ep[qjm]);
            const T eyv = gs05(-Zf[FNQ + q], -Zf[FNQ + qim], fxbz_c, fxbz_p, Zf[q], Zf[qim], mfx_c, mfx_p,
                               Ec[FNQ + q], Ep[FNQ + q], Ec[FNQ + qim], Ep[FNQ + qim]);
            const int64_t idx = (int64_t)m * g.plane + col;
            ctslot(15, idx) = exv;
            ctslot(16, idx) = eyv;
            ctslot(17, idx) = ez;
          }
          fybz_p = fybz_c;
          mfy_p = mfy_c;
          fxbz_p = fxbz_c;
          mfx_p = mfx_c;
        }
      }
    }
  }
  return smax;
}

// Block-wide CFL maximum of the threads' running maxima, stored by thread 0.
template <class T>
__device__ __forceinline__ void block_store_max(T smax, T* sMax, double* out) {
  const int t = threadIdx.x;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) smax = nanmax(smax, __shfl_xor(smax, o, 64));
  if ((t & 63) == 0) sMax[t >> 6] = smax;
  __syncthreads();
  if (t == 0) {
    T r = sMax[0];
    for (int i = 1; i < FTHREADS / 64; ++i) r = nanmax(r, sMax[i]);
    *out = (double)r;
  }
}

// Three waves per SIMD (168 VGPRs; three 38 KB workgroups then share a CU's LDS) for
// every plain float variant and the PLM source variants. PLM fits the plain step without
// spilling; PPM keeps both edge offsets of each direction live and is still faster at
// three waves than at two. The PLM source variants (drag, radiation) spill 4-7 dwords at
// three waves and run 14-17% faster than at two; the PPM source variants would spill 44
// and keep the compiler's default (two waves), as do the surrogate and float64 variants.
template <class T, int RECON, int RS, bool SURR, bool SRC>
constexpr bool fused_three_waves() {
  return sizeof(T) == 4 && !SURR && !(SRC && RECON == RECON_PPM);
}

// The single-grid step: one workgroup per work item. simg is the network's image in
// global memory (SURR only) and nsmooth the count of network-solved faces.
template <class T, int RECON, int RS, bool SURR, bool SRC, bool NARROW, bool BCG = false>
__device__ __forceinline__ void step_fused_body(const Fields<const T>& U, const Fields<T>& O, const FusedGeom& g,
                                                const StepParams<T>& P, double* __restrict__ red,
                                                const unsigned char* __restrict__ simg, float thr,
                                                unsigned long long* __restrict__ nsmooth, const SrcParams<T>& src) {
  __shared__ FusedLds<T, SRC, dxy_in_lds<RECON, SURR>()> lds;
  __shared__ T sMax[FTHREADS / 64];
  __shared__ __attribute__((aligned(16))) unsigned short sW[SURR ? SIMG_NW : 2];
  __shared__ __attribute__((aligned(16))) SurrQueue<SURR ? FQCAP : 1> sQ;
  unsigned smooth_count = 0;
  if constexpr (SURR) {
    const unsigned* src = reinterpret_cast<const unsigned*>(simg);
    unsigned* dw = reinterpret_cast<unsigned*>(sW);
    for (int e = threadIdx.x; e < SIMG_NW / 2; e += FTHREADS) dw[e] = src[e];
  }
  // XCD-aware tile order: dispatch round-robins workgroups over the 8 XCDs, so
  // logical tile (w % 8) * (nwg / 8) + w / 8 gives each XCD a contiguous tile range.
  const int nwg = gridDim.x;
  int w = blockIdx.x;
  if ((nwg & 7) == 0) w = (w & 7) * (nwg >> 3) + (w >> 3);
  SurrQueue<FQCAP>* q = nullptr;
  if constexpr (SURR) q = &sQ;
  // the biases (the image's float tail) stay in global memory: with them out of LDS the
  // surrogate step's 81.6 KB lets two workgroups share a CU
  const float* sB = SURR ? reinterpret_cast<const float*>(simg + SIMG_NW * 2) : nullptr;
  const T smax =
      fused_item<T, RECON, RS, SURR, SRC, NARROW, false, BCG>(U, O, g, P, w, lds, sW, sB, thr, smooth_count, src, q);
  // fold the workgroup's CFL maximum (timestep.jl:8-12 on the new state)
  block_store_max(smax, sMax, red + blockIdx.x);
  if constexpr (SURR) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) smooth_count += __shfl_xor(smooth_count, o, 64);
    if ((threadIdx.x & 63) == 0 && smooth_count) atomicAdd(nsmooth, (unsigned long long)smooth_count);
  }
}

// Pd: when not null, the step's scalars come from device memory (written there by the
// device-side CFL stage of pmf_advance) instead of the kernel argument P.
template <class T, int RECON, int RS, bool SURR, bool SRC, bool NARROW, bool BCG>
__global__ __launch_bounds__(FTHREADS) void k_step_fused(Fields<const T> U, Fields<T> O, FusedGeom g,
                                                         StepParams<T> P, const StepParams<T>* __restrict__ Pd,
                                                         double* __restrict__ red,
                                                         const unsigned char* __restrict__ simg, float thr,
                                                         unsigned long long* __restrict__ nsmooth, SrcParams<T> src) {
  if (Pd) P = *Pd;
  step_fused_body<T, RECON, RS, SURR, SRC, NARROW, BCG>(U, O, g, P, red, simg, thr, nsmooth, src);
}

template <class T, int RECON, int RS, bool SURR, bool SRC, bool NARROW, bool BCG>
__global__ __launch_bounds__(FTHREADS) __attribute__((amdgpu_waves_per_eu(3))) void k_step_fused3(
    Fields<const T> U, Fields<T> O, FusedGeom g, StepParams<T> P, const StepParams<T>* __restrict__ Pd,
    double* __restrict__ red, const unsigned char* __restrict__ simg, float thr,
    unsigned long long* __restrict__ nsmooth, SrcParams<T> src) {
  if (Pd) P = *Pd;
  step_fused_body<T, RECON, RS, SURR, SRC, NARROW, BCG>(U, O, g, P, red, simg, thr, nsmooth, src);
}

// The surrogate-assisted step: two waves per SIMD (about 250 VGPRs with the network in its
// own function; 81.6 KB of LDS), so two workgroups share a CU and one's network batches
// overlap the other's HLLD faces and barriers.
template <class T, int RECON, int RS, bool SURR, bool SRC, bool NARROW>
__global__ __launch_bounds__(FTHREADS) __attribute__((amdgpu_waves_per_eu(2))) void k_step_fused2(
    Fields<const T> U, Fields<T> O, FusedGeom g, StepParams<T> P, double* __restrict__ red,
    const unsigned char* __restrict__ simg, float thr, unsigned long long* __restrict__ nsmooth, SrcParams<T> src) {
  step_fused_body<T, RECON, RS, SURR, SRC, NARROW>(U, O, g, P, red, simg, thr, nsmooth, src);
}

template <class T, int RECON, int RS, bool SURR, bool SRC = false, bool NARROW = false, bool BCG = false>
void launch_fused_variant(dim3 grd, hipStream_t st, const Fields<const T>& U, const Fields<T>& O, const FusedGeom& g,
                          const StepParams<T>& P, double* red, const unsigned char* simg, float thr,
                          unsigned long long* nsmooth, const SrcParams<T>& src = SrcParams<T>{},
                          const StepParams<T>* Pd = nullptr) {
  if constexpr (fused_three_waves<T, RECON, RS, SURR, SRC>())
    hipLaunchKernelGGL((k_step_fused3<T, RECON, RS, SURR, SRC, NARROW, BCG>), grd, dim3(FTHREADS), 0, st, U, O, g, P,
                       Pd, red, simg, thr, nsmooth, src);
  else if constexpr (SURR)
    hipLaunchKernelGGL((k_step_fused2<T, RECON, RS, SURR, SRC, NARROW>), grd, dim3(FTHREADS), 0, st, U, O, g, P, red,
                       simg, thr, nsmooth, src);
  else
    hipLaunchKernelGGL((k_step_fused<T, RECON, RS, SURR, SRC, NARROW, BCG>), grd, dim3(FTHREADS), 0, st, U, O, g, P,
                       Pd, red, simg, thr, nsmooth, src);
}

// The plain float step with 32-bit offsets (double steps never take this path).
template <class T, int RECON, int RS>
void launch_plain_narrow(dim3 grd, hipStream_t st, const Fields<const T>& U, const Fields<T>& O, const FusedGeom& g,
                         const StepParams<T>& P, double* red, const StepParams<T>* Pd) {
  if constexpr (sizeof(T) == 4)
    launch_fused_variant<T, RECON, RS, false, false, true>(grd, st, U, O, g, P, red, nullptr, 0.f, nullptr,
                                                           SrcParams<T>{}, Pd);
}

// Fields of nc cells addressable with 32-bit byte offsets. PMFMHD_WIDE_OFFSETS=1 in the
// environment selects the 64-bit-offset kernels for every grid (they otherwise run only
// for float fields of 4 GiB and more, 1024^3 and up), so that tests can compare the two.
template <class T> inline bool narrow_fields(int64_t nc) {
  const char* wide = std::getenv("PMFMHD_WIDE_OFFSETS");
  if (wide && wide[0] == '1') return false;
  return (uint64_t)nc * sizeof(T) < (1ull << 32);
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

// The device-side CFL stage of pmf_advance: evolve!'s step rule (driver.jl:48-71) from
// the speed the previous step folded into dsmax[0], in the same operations and
// precisions as the host path (pmf_compute_dt, pmf_evolve, decay_of, make_params), so
// that a step with device-resident scalars is the host-driven step bit for bit.
template <class T>
__global__ __launch_bounds__(64) void k_step_params(const double* __restrict__ dsmax, StepParams<T> base,
                                                    StepParams<T>* __restrict__ Pd, double* __restrict__ clock,
                                                    double cfl, double fac, double cr, double dx, int reset) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  const double sm = dsmax[0];
  double dtf, ch, decay;
  if constexpr (sizeof(T) == 8) {
    const double dt = cfl * dx / (sm < __DBL_EPSILON__ ? __DBL_EPSILON__ : sm);
    ch = fac * sm;
    dtf = dt * sm / (sm < ch ? ch : sm);
    decay = exp(-cr * ch * dtf / dx);
  } else {
    const float smf = float(sm), dxf = float(dx);
    const double dt = double(float(cfl) * dxf / (smf < __FLT_EPSILON__ ? __FLT_EPSILON__ : smf));
    const float chf = float(fac) * smf;
    ch = chf;
    dtf = dt * double(smf) / double(smf < chf ? chf : smf);
    const float v = -float(cr) * chf * float(dtf) / dxf;
    decay = double(float(exp(double(v))));
  }
  StepParams<T> P = base;
  P.ch = T(ch);
  P.ich = ch != 0.0 ? T(1) / T(ch) : T(0);
  P.ch2 = T(ch) * T(ch);
  P.dtdx = T(dtf) / T(dx);
  P.hdtdx = T(0.5) * P.dtdx;
  P.decay = T(decay);
  *Pd = P;
  clock[0] = reset ? dtf : clock[0] + dtf;
  clock[1] = dtf;
}

template <class T>
void launch_step_params(const double* dsmax, const StepParams<T>& base, StepParams<T>* Pd, double* clock, double cfl,
                        double glm_ch_fac, double glm_cr, double dx, int reset, hipStream_t st) {
  hipLaunchKernelGGL((k_step_params<T>), dim3(1), dim3(64), 0, st, dsmax, base, Pd, clock, cfl, glm_ch_fac, glm_cr, dx,
                     reset);
}
template void launch_step_params<float>(const double*, const StepParams<float>&, StepParams<float>*, double*, double,
                                        double, double, double, int, hipStream_t);
template void launch_step_params<double>(const double*, const StepParams<double>&, StepParams<double>*, double*,
                                         double, double, double, double, int, hipStream_t);

static FusedGeom fused_geom(const Grid& G, int zc) {
  FusedGeom g;
  g.nx = G.nx;
  g.ny = G.ny;
  g.nz = G.nz;
  g.ntx = (G.nx + TX - 1) / TX;
  g.nty = (G.ny + TY - 1) / TY;
  g.zc = zc;
  g.nchunks = (G.nz + zc - 1) / zc;
  g.plane = (int64_t)G.nx * G.ny;
  g.nc = G.nc;
  for (int d = 0; d < 3; ++d) g.bc[d] = G.bc[d];
  return g;
}

// Any grid: ragged last tiles and chunks are computed and not stored, and out-of-range
// columns and planes come from the boundary map. A reflecting axis needs two cells (its
// second ghost layer mirrors onto the second cell, boundary.jl:24-28).
bool fused_supported(const Grid& G, int precision, int recon, int rsolve) {
  (void)precision;
  (void)recon;
  (void)rsolve;
  const int32_t n[3] = {G.nx, G.ny, G.nz};
  for (int d = 0; d < 3; ++d)
    if (n[d] < 1 || (G.bc[d] == BC_REFLECT && n[d] < 2)) return false;
  return true;
}

bool all_periodic(const Grid& G) {
  return G.bc[0] == BC_PERIODIC && G.bc[1] == BC_PERIODIC && G.bc[2] == BC_PERIODIC;
}

// The surrogate step's grids: periodic, whole x tiles, ny and nz of at least two.
bool surrogate_grid(const Grid& G) { return all_periodic(G) && G.nx % TX == 0 && G.ny >= 2 && G.nz >= 2; }

// Compute units of the current device (cached per device).
static int device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cus[dev]) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cus[dev] = n;
  }
  return cus[dev];
}

// z-chunk length: long enough that the two extra prediction planes per chunk stay
// small, short enough to give a few workgroups per CU. Halving from 32 until the grid
// has at least 2048 workgroups suits large grids; on small ones it ends at 4-8 planes,
// where a chunk of 16 whose workgroups fill most of one round (three per CU) wastes
// fewer planes (128^3: 704 workgroups, 0.136 vs 0.153 ms at 4; at 160^3, 96^3 and 192^3
// the halving rule's choice measured best).
static int pick_zc(const Grid& G) {
  const int64_t tiles = (int64_t)((G.nx + TX - 1) / TX) * ((G.ny + TY - 1) / TY);
  auto groups = [&](int zc) { return tiles * ((G.nz + zc - 1) / zc); };
  int zc = 32;
  while (zc > 4 && groups(zc) < 2048) zc >>= 1;
  const int64_t round = 3 * (int64_t)device_cus();
  if (zc <= 8 && groups(16) <= round && 3 * groups(16) >= 2 * round) zc = 16;
  return zc < G.nz ? zc : G.nz;
}

static int chunk_of(const Grid& G, int zc) { return zc > 0 ? (zc < G.nz ? zc : G.nz) : pick_zc(G); }

int64_t fused_blocks(const Grid& G, int zc) {
  FusedGeom g = fused_geom(G, chunk_of(G, zc));
  return (int64_t)g.ntx * g.nty * g.nchunks;
}

template <class T>
bool launch_step_fused(Fields<const T> U, Fields<T> O, const Grid& G, const StepParams<T>& P, double* red,
                       double* dsmax, hipStream_t st, int zc, const SrcParams<T>* src, const StepParams<T>* Pd) {
  if (!fused_supported(G, sizeof(T) == 8, P.recon, P.rsolve)) return false;
  FusedGeom g = fused_geom(G, chunk_of(G, zc));
  unsigned nb = (unsigned)((int64_t)g.ntx * g.nty * g.nchunks);
  dim3 blk(FTHREADS), grd(nb);
  const bool narrow = sizeof(T) == 4 && narrow_fields<T>(G.nc);
  const bool gen = !all_periodic(G);  // outflow / reflecting axes: the BCG variants
#define PMF_FUSED(R, S)                                                                                             \
  if (gen && src)                                                                                                  \
    launch_fused_variant<T, R, S, false, true, false, true>(grd, st, U, O, g, P, red, nullptr, 0.f, nullptr, *src); \
  else if (gen) launch_fused_variant<T, R, S, false, false, false, true>(grd, st, U, O, g, P, red, nullptr, 0.f,     \
                                                                         nullptr, SrcParams<T>{}, Pd);              \
  else if (src) launch_fused_variant<T, R, S, false, true>(grd, st, U, O, g, P, red, nullptr, 0.f, nullptr, *src);   \
  else if (narrow) launch_plain_narrow<T, R, S>(grd, st, U, O, g, P, red, Pd);                                      \
  else launch_fused_variant<T, R, S, false, false>(grd, st, U, O, g, P, red, nullptr, 0.f, nullptr, SrcParams<T>{}, Pd)
#define PMF_FUSED_RS(R)                                  \
  if (P.rsolve == RS_HLL) PMF_FUSED(R, RS_HLL);          \
  else if (P.rsolve == RS_LLF) PMF_FUSED(R, RS_LLF);     \
  else PMF_FUSED(R, RS_HLLD);
  if (P.recon == RECON_PPM) {
    PMF_FUSED_RS(RECON_PPM)
  } else if (P.recon == RECON_PLM_MINMOD) {
    PMF_FUSED_RS(RECON_PLM_MINMOD)
  } else {
    PMF_FUSED_RS(RECON_PLM)
  }
#undef PMF_FUSED_RS
#undef PMF_FUSED
  hipLaunchKernelGGL(k_fold_max_flat, dim3(1), dim3(256), 0, st, red, (int)nb, dsmax);
  return true;
}

bool launch_step_fused_surr(Fields<const float> U, Fields<float> O, const Grid& G, const StepParams<float>& P,
                            double* red, double* dsmax, const unsigned char* simg, float thr,
                            unsigned long long* nsmooth, hipStream_t st) {
  if (!surrogate_grid(G) || P.rsolve != RS_HLLD) return false;
  FusedGeom g = fused_geom(G, pick_zc(G));
  unsigned nb = (unsigned)((int64_t)g.ntx * g.nty * g.nchunks);
  dim3 blk(FTHREADS), grd(nb);
#define PMF_FUSED_S(R) launch_fused_variant<float, R, RS_HLLD, true>(grd, st, U, O, g, P, red, simg, thr, nsmooth)
  if (P.recon == RECON_PPM) PMF_FUSED_S(RECON_PPM);
  else if (P.recon == RECON_PLM_MINMOD) PMF_FUSED_S(RECON_PLM_MINMOD);
  else PMF_FUSED_S(RECON_PLM);
#undef PMF_FUSED_S
  hipLaunchKernelGGL(k_fold_max_flat, dim3(1), dim3(256), 0, st, red, (int)nb, dsmax);
  return true;
}

// ---------------------------------------------------------------------------
// Ensemble: many independent periodic cubes advanced together. Each stage is one
// launch over all cubes: the step is a persistent kernel whose workgroups draw work
// items (cube, tile) from eight queues, one per XCD, each a contiguous cost-balanced
// slice of the item list (so neighbouring tiles share an L2), stealing from the other
// queues once their own is empty; the per-cube time step comes from a device-side
// stage that folds the step's per-item CFL maxima. No host round trip per step, so
// the step pair can be captured in a graph.

__global__ __launch_bounds__(FTHREADS) void k_ens_wavespeed(const EnsCube* __restrict__ cubes,
                                                          const int2* __restrict__ items, StepParams<float> P,
                                                          double* __restrict__ red, int parity) {
  __shared__ float sMax[FTHREADS / 64];
  const int2 it = items[blockIdx.x];
  const int c = __builtin_amdgcn_readfirstlane(it.x), w = __builtin_amdgcn_readfirstlane(it.y);
  const EnsCube& cb = cubes[c];
  const FusedGeom& g = cb.g;
  const Fields<float>& U = parity ? cb.B : cb.A;
  const int tx = w % g.ntx, ty = (w / g.ntx) % g.nty, chunk = w / (g.ntx * g.nty);
  const int z0 = chunk * g.zc, zc = min(g.zc, g.nz - z0);
  const int t = threadIdx.x;
  float smax = 0.f;
  if (t < NOWN) {
    const int ix = tx * TX + t % TX, iy = ty * TY + t / TX;
    if (iy < g.ny) {
      for (int z = z0; z < z0 + zc; ++z) {
        const int64_t idx = (int64_t)z * g.plane + (int64_t)iy * g.nx + ix;
        float q[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) q[v] = U.v[v][idx];
        smax = nanmax(smax, cell_wavespeed(q, P));
      }
    }
  }
  block_store_max(smax, sMax, red + blockIdx.x);
}

// Per-cube CFL stage (timestep.jl:39 and driver.jl:48-71 in float32 arithmetic, the same
// operations as the single-state host path): fold the cube's item maxima, then
// dt = cfl dx / max(smax, eps), ch = glm_ch_fac smax, dt_used = dt smax / max(smax, ch),
// decay = exp(-cr ch dt_used / dx). Also clears the step's work queues.
__global__ __launch_bounds__(256) void k_ens_params(const EnsCube* __restrict__ cubes, const double* __restrict__ red,
                                                    StepParams<float> base, StepParams<float>* __restrict__ params,
                                                    double* __restrict__ clock, unsigned* __restrict__ ctr, float cfl,
                                                    float fac, float cr) {
  __shared__ double sh[4];
  const int c = blockIdx.x;
  const EnsCube& cb = cubes[c];
  double m = 0.0;
  for (int i = threadIdx.x; i < cb.nitems; i += 256) m = nanmax(m, red[cb.item0 + i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = nanmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double smax = nanmax(nanmax(sh[0], sh[1]), nanmax(sh[2], sh[3]));
    const float smf = float(smax), dxf = float(cb.dx);
    const double dt = double(cfl * dxf / fmaxf(smf, __FLT_EPSILON__));
    const float chf = fac * smf;
    const double dtf = dt * double(smf) / double(fmaxf(smf, chf));
    StepParams<float> P = base;
    P.ch = chf;
    P.ich = chf != 0.f ? 1.f / chf : 0.f;
    P.ch2 = chf * chf;
    P.dtdx = float(dtf) / dxf;
    P.hdtdx = 0.5f * P.dtdx;
    const float v = -cr * chf * float(dtf) / dxf;
    P.decay = float(exp(double(v)));
    params[c] = P;
    clock[3 * c] += dtf;
    clock[3 * c + 1] = dtf;
    clock[3 * c + 2] = smax;
    if (c == 0) {
      for (int q = 0; q < 8; ++q) ctr[q] = 0u;
    }
  }
}

template <int RECON, int RS>
__device__ __forceinline__ void ens_step_body(const EnsCube* __restrict__ cubes, const int2* __restrict__ items,
                                              const EnsQueues& q, const StepParams<float>* __restrict__ params,
                                              double* __restrict__ red, unsigned* __restrict__ ctr, int parity) {
  __shared__ FusedLds<float, false, dxy_in_lds<RECON, false>()> lds;
  __shared__ float sMax[FTHREADS / 64];
  __shared__ int sItem;
  unsigned none = 0;
  int qi = 0;  // queues tried so far (thread 0)
  const int home = blockIdx.x & 7;
  for (;;) {
    if (threadIdx.x == 0) {
      int got = -1;
      while (qi < 8) {
        const int qq = (home + qi) & 7;
        const unsigned k = atomicAdd(&ctr[qq], 1u);
        if (k < (unsigned)q.len[qq]) {
          got = q.base[qq] + (int)k;
          break;
        }
        ++qi;
      }
      sItem = got;
    }
    __syncthreads();
    const int i = __builtin_amdgcn_readfirstlane(sItem);
    if (i < 0) break;  // uniform: every queue is drained
    const int2 it = items[i];
    const int c = __builtin_amdgcn_readfirstlane(it.x), w = __builtin_amdgcn_readfirstlane(it.y);
    const EnsCube& cb = cubes[c];
    const Fields<float>& in = parity ? cb.B : cb.A;
    const Fields<float>& out = parity ? cb.A : cb.B;
    Fields<const float> U;
#pragma unroll
    for (int v = 0; v < 9; ++v) U.v[v] = in.v[v];
    const StepParams<float> P = params[c];
    const float smax =
        fused_item<float, RECON, RS, false, false, true>(U, out, cb.g, P, w, lds, nullptr, nullptr, 0.f, none,
                                                         SrcParams<float>{});
    block_store_max(smax, sMax, red + i);
    __syncthreads();  // sItem and sMax are reused by the next item
  }
}

template <int RECON, int RS>
__global__ __launch_bounds__(FTHREADS) void k_ens_step(const EnsCube* __restrict__ cubes,
                                                     const int2* __restrict__ items, EnsQueues q,
                                                     const StepParams<float>* __restrict__ params,
                                                     double* __restrict__ red, unsigned* __restrict__ ctr,
                                                     int parity) {
  ens_step_body<RECON, RS>(cubes, items, q, params, red, ctr, parity);
}

template <int RECON, int RS>
__global__ __launch_bounds__(FTHREADS) __attribute__((amdgpu_waves_per_eu(3))) void k_ens_step3(
    const EnsCube* __restrict__ cubes, const int2* __restrict__ items, EnsQueues q,
    const StepParams<float>* __restrict__ params, double* __restrict__ red, unsigned* __restrict__ ctr, int parity) {
  ens_step_body<RECON, RS>(cubes, items, q, params, red, ctr, parity);
}

FusedGeom ens_geom(int n, int zc) {
  Grid G;
  G.nx = G.ny = G.nz = n;
  G.bc[0] = G.bc[1] = G.bc[2] = BC_PERIODIC;
  G.nc = (int64_t)n * n * n;
  return fused_geom(G, zc);
}

void launch_ens_wavespeed(const EnsCube* cubes, const int2* items, int nitems, const StepParams<float>& P,
                          double* red, int parity, hipStream_t st) {
  hipLaunchKernelGGL(k_ens_wavespeed, dim3(nitems), dim3(FTHREADS), 0, st, cubes, items, P, red, parity);
}

void launch_ens_params(const EnsCube* cubes, int ncubes, const double* red, const StepParams<float>& base,
                       StepParams<float>* params, double* clock, unsigned* ctr, float cfl, float fac, float cr,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_ens_params, dim3(ncubes), dim3(256), 0, st, cubes, red, base, params, clock, ctr, cfl, fac,
                     cr);
}

void launch_ens_step(const EnsCube* cubes, const int2* items, const EnsQueues& q, const StepParams<float>* params,
                     double* red, unsigned* ctr, int nwg, int parity, int recon, int rsolve, hipStream_t st) {
#define PMF_ENS(R, S)                                                                                     \
  if constexpr (fused_three_waves<float, R, S, false, false>())                                                     \
    hipLaunchKernelGGL((k_ens_step3<R, S>), dim3(nwg), dim3(FTHREADS), 0, st, cubes, items, q, params, red, ctr, \
                       parity);                                                                              \
  else                                                                                                       \
    hipLaunchKernelGGL((k_ens_step<R, S>), dim3(nwg), dim3(FTHREADS), 0, st, cubes, items, q, params, red, ctr, parity)
#define PMF_ENS_RS(R)                                 \
  if (rsolve == RS_HLL) PMF_ENS(R, RS_HLL);           \
  else if (rsolve == RS_LLF) PMF_ENS(R, RS_LLF);      \
  else PMF_ENS(R, RS_HLLD);
  if (recon == RECON_PPM) {
    PMF_ENS_RS(RECON_PPM)
  } else if (recon == RECON_PLM_MINMOD) {
    PMF_ENS_RS(RECON_PLM_MINMOD)
  } else {
    PMF_ENS_RS(RECON_PLM)
  }
#undef PMF_ENS_RS
#undef PMF_ENS
}

// The fused CT step: the :ref algorithm for constrained-transport states (riemann_ct on
// the face fields, integrator_ref.jl:83-139) with each face solved once; writes the cell
// update and workspace slots 3-14 (face mass and transverse-B fluxes of every cell's
// lower faces, cell EMFs) that ct_update! reads.
template <class T, int RECON, int RS, bool SRC = false, bool NARROW = false, bool EDGES = false>
__device__ __forceinline__ void step_fused_ct_body(const Fields<const T>& U, const Fields<T>& O, T* __restrict__ ctp,
                                                   const FusedGeom& g, const StepParams<T>& P,
                                                   const SrcParams<T>& src = SrcParams<T>{}) {
  __shared__ FusedLds<T, SRC, dxy_in_lds<RECON, false>(), EDGES> lds;
  const int nwg = gridDim.x;
  int w = blockIdx.x;
  if ((nwg & 7) == 0) w = (w & 7) * (nwg >> 3) + (w >> 3);
  unsigned none = 0;
  (void)fused_item<T, RECON, RS, false, SRC, NARROW, true, false, EDGES>(U, O, g, P, w, lds, nullptr, nullptr, 0.f,
                                                                         none, src, nullptr, ctp);
}

template <class T, int RECON, int RS>
__global__ __launch_bounds__(FTHREADS) void k_step_fused_ct(Fields<const T> U, Fields<T> O, T* __restrict__ ctp,
                                                            FusedGeom g, StepParams<T> P) {
  step_fused_ct_body<T, RECON, RS>(U, O, ctp, g, P);
}

// float: three waves per SIMD, as the plain step, where that does not spill (PPM would
// otherwise take 174 VGPRs and run two waves: 1.81 vs 2.08 ms per 256^3 CT step with
// HLLD). minmod-LLF spills at three waves and keeps two.
template <int RECON, int RS>
constexpr bool ct_three_waves() {
  return !(RECON == RECON_PLM_MINMOD && RS == RS_LLF);
}

// NARROW: every field and every slot of the packed buffer is addressed as its scalar base
// plus a 32-bit byte offset (fields below 4 GiB). EDGES: the step forms the edge EMFs
// (slots 15-17) itself and leaves slots 3-14 alone.
template <int RECON, int RS, bool NARROW, bool EDGES = false>
__global__ __launch_bounds__(FTHREADS) __attribute__((amdgpu_waves_per_eu(3))) void k_step_fused_ct3(
    Fields<const float> U, Fields<float> O, float* __restrict__ ctp, FusedGeom g, StepParams<float> P) {
  step_fused_ct_body<float, RECON, RS, false, NARROW, EDGES>(U, O, ctp, g, P);
}

// The in-step edge EMFs: float, three waves, no sources, and PLM (PPM's LDS divergence
// carry leaves no room for the exchange arrays at three workgroups per CU).
template <int RECON, int RS>
constexpr bool ct_edges_in_step() {
  return RECON != RECON_PPM && ct_three_waves<RECON, RS>();
}

// The CT step with the drag / radiation sources (step_ref_drag! / step_ref_radiation! on
// CT states, integrator_ref.jl:222-258): source-modified face states and cell EMFs,
// tracked updates (no cell-B tracking under CT), the admissibility mark.
template <class T, int RECON, int RS>
__global__ __launch_bounds__(FTHREADS) void k_step_fused_ct_src(Fields<const T> U, Fields<T> O, T* __restrict__ ctp,
                                                                FusedGeom g, StepParams<T> P, SrcParams<T> src) {
  step_fused_ct_body<T, RECON, RS, true>(U, O, ctp, g, P, src);
}

template <class T>
bool launch_step_fused_ct(Fields<const T> U, Fields<T> O, T* ctp, const Grid& G, const StepParams<T>& P,
                          hipStream_t st, int zc, const SrcParams<T>* src, bool want_edges, bool* edges) {
  if (edges) *edges = false;
  if (!fused_supported(G, sizeof(T) == 8, P.recon, P.rsolve) || !all_periodic(G)) return false;
  FusedGeom g = fused_geom(G, chunk_of(G, zc));
  dim3 blk(FTHREADS), grd((unsigned)((int64_t)g.ntx * g.nty * g.nchunks));
  const bool narrow = sizeof(T) == 4 && narrow_fields<T>(G.nc);
#define PMF_CT(R, S)                                                                   \
  do {                                                                                 \
    if (src)                                                                           \
      hipLaunchKernelGGL((k_step_fused_ct_src<T, R, S>), grd, blk, 0, st, U, O, ctp, g, P, *src); \
    else if constexpr (sizeof(T) == 4 && ct_three_waves<R, S>()) {                    \
      if constexpr (ct_edges_in_step<R, S>()) {                                        \
        if (want_edges) {                                                              \
          if (narrow)                                                                  \
            hipLaunchKernelGGL((k_step_fused_ct3<R, S, true, true>), grd, blk, 0, st, U, O, ctp, g, P); \
          else                                                                         \
            hipLaunchKernelGGL((k_step_fused_ct3<R, S, false, true>), grd, blk, 0, st, U, O, ctp, g, P); \
          if (edges) *edges = true;                                                    \
          break;                                                                       \
        }                                                                              \
      }                                                                                \
      if (narrow)                                                                      \
        hipLaunchKernelGGL((k_step_fused_ct3<R, S, true>), grd, blk, 0, st, U, O, ctp, g, P); \
      else                                                                             \
        hipLaunchKernelGGL((k_step_fused_ct3<R, S, false>), grd, blk, 0, st, U, O, ctp, g, P); \
    } else                                                                             \
      hipLaunchKernelGGL((k_step_fused_ct<T, R, S>), grd, blk, 0, st, U, O, ctp, g, P); \
  } while (0)
#define PMF_CT_RS(R)                                \
  if (P.rsolve == RS_HLL) PMF_CT(R, RS_HLL);        \
  else if (P.rsolve == RS_LLF) PMF_CT(R, RS_LLF);   \
  else PMF_CT(R, RS_HLLD);
  if (P.recon == RECON_PPM) {
    PMF_CT_RS(RECON_PPM)
  } else if (P.recon == RECON_PLM_MINMOD) {
    PMF_CT_RS(RECON_PLM_MINMOD)
  } else {
    PMF_CT_RS(RECON_PLM)
  }
#undef PMF_CT_RS
#undef PMF_CT
  return true;
}

template bool launch_step_fused_ct<float>(Fields<const float>, Fields<float>, float*, const Grid&,
                                          const StepParams<float>&, hipStream_t, int, const SrcParams<float>*, bool,
                                          bool*);
template bool launch_step_fused_ct<double>(Fields<const double>, Fields<double>, double*, const Grid&,
                                           const StepParams<double>&, hipStream_t, int, const SrcParams<double>*, bool,
                                           bool*);

template bool launch_step_fused<float>(Fields<const float>, Fields<float>, const Grid&, const StepParams<float>&,
                                       double*, double*, hipStream_t, int, const SrcParams<float>*,
                                       const StepParams<float>*);
template bool launch_step_fused<double>(Fields<const double>, Fields<double>, const Grid&, const StepParams<double>&,
                                        double*, double*, hipStream_t, int, const SrcParams<double>*,
                                        const StepParams<double>*);

}  // namespace pmf
