// Host-side launchers of the HIP kernels (implemented in the .hip files).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mhd_device.hpp"
#include "sources.hpp"

namespace pmf {

// Size of the per-block partial-result workspace (doubles) that the reductions use.
constexpr int kRedBlocks = 2048;
constexpr int kRedSlots = 24;  // doubles per block slot

// S: drag / radiation source terms (nullptr for the plain step).
template <class T>
void launch_step_ref(Fields<const T> U, Fields<T> O, T* ctp, const Grid& G, const StepParams<T>& P, bool ct,
                     hipStream_t st, const SrcParams<T>* S = nullptr);

// Fused tiled step (periodic GLM). Writes per-block CFL maxima of the new state into
// red (one double per block) and the final max into dsmax. Returns false when the grid
// shape is not supported by the fused kernel.
// zc: planes per z chunk (0 = the built-in choice); src: drag / radiation source terms
// (nullptr for the plain step).
// Pd: device-resident step scalars that replace P's (dt, ch, decay) in the kernel (P still
// selects the reconstruction and Riemann solver); nullptr = P as passed.
template <class T>
bool launch_step_fused(Fields<const T> U, Fields<T> O, const Grid& G, const StepParams<T>& P, double* red,
                       double* dsmax, hipStream_t st, int zc = 0, const SrcParams<T>* src = nullptr,
                       const StepParams<T>* Pd = nullptr);
// The device-side CFL stage of pmf_advance (evolve!'s rule, driver.jl:48-71, evaluated as
// the host does it): from the speed in dsmax[0], writes the step scalars to Pd and adds
// the step to clock[0] (reset: sets it), clock[1] = the step's dt.
template <class T>
void launch_step_params(const double* dsmax, const StepParams<T>& base, StepParams<T>* Pd, double* clock, double cfl,
                        double glm_ch_fac, double glm_cr, double dx, int reset, hipStream_t st);
bool fused_supported(const Grid& G, int precision, int recon, int rsolve);
bool all_periodic(const Grid& G);
bool surrogate_grid(const Grid& G);
// The fused CT step (constrained-transport states, the :ref algorithm with each face
// solved once): the cell update into O and workspace slots 3-14 of ctp; ct_update! then
// finishes the step. src: the drag / radiation sources (no fallback masks: those belong
// to the :cube integrator). want_edges: where the variant supports it, form the edge
// EMFs (slots 15-17) in the step instead of writing slots 3-14; *edges tells whether it
// did (then ct_update! skips its edge pass). Returns false when the grid shape is not
// supported.
template <class T>
bool launch_step_fused_ct(Fields<const T> U, Fields<T> O, T* ctp, const Grid& G, const StepParams<T>& P,
                          hipStream_t st, int zc = 0, const SrcParams<T>* src = nullptr, bool want_edges = false,
                          bool* edges = nullptr);
// The surrogate-assisted fused step (float32, HLLD): simg is the network's LDS image,
// nsmooth accumulates the number of faces solved by the network.
bool launch_step_fused_surr(Fields<const float> U, Fields<float> O, const Grid& G, const StepParams<float>& P,
                            double* red, double* dsmax, const unsigned char* simg, float thr,
                            unsigned long long* nsmooth, hipStream_t st);
// Number of workgroups (one partial maximum each) the fused kernel launches for G with
// z chunks of zc planes (0 = the built-in choice).
int64_t fused_blocks(const Grid& G, int zc = 0);
constexpr int kMinChunk = 4;  // smallest z chunk the library accepts

// Geometry of the fused kernel's work decomposition: 16x12 column tiles (x fastest,
// then y; the last tile row may be ragged) marched through z chunks of zc planes.
struct FusedGeom {
  int32_t nx, ny, nz, ntx, nty, nchunks, zc;
  int32_t bc[3];  // per-axis boundary codes (BC_PERIODIC, BC_OUTFLOW, BC_REFLECT)
  int64_t plane, nc;
};
constexpr int kFusedTX = 16, kFusedTY = 12;

// One cube of an ensemble: both field sets (the step reads A and writes B on even
// steps, the reverse on odd ones), its geometry and its slice of the work-item list.
struct EnsCube {
  Fields<float> A, B;
  FusedGeom g;
  double dx;
  int32_t item0, nitems;
};
// Eight work queues (one per XCD): contiguous slices [base, base + len) of the items.
struct EnsQueues {
  int32_t base[8], len[8];
};
FusedGeom ens_geom(int n, int zc);
// Initial CFL maxima, one per work item (items are (cube, tile) pairs).
void launch_ens_wavespeed(const EnsCube* cubes, const int2* items, int nitems, const StepParams<float>& P,
                          double* red, int parity, hipStream_t st);
// Per-cube dt / ch / decay from the item maxima in red; clock holds (t, dt, smax) per cube.
void launch_ens_params(const EnsCube* cubes, int ncubes, const double* red, const StepParams<float>& base,
                       StepParams<float>* params, double* clock, unsigned* ctr, float cfl, float fac, float cr,
                       hipStream_t st);
// The persistent step over all items (nwg workgroups; ctr = eight queue counters).
void launch_ens_step(const EnsCube* cubes, const int2* items, const EnsQueues& q, const StepParams<float>* params,
                     double* red, unsigned* ctr, int nwg, int parity, int recon, int rsolve, hipStream_t st);

// Passive species (passive_transport.jl): fractions from conserved densities in place
// (s2 may be null), and the backtrace into o1 (o2) of fractions (mode 0 / 1) or of an
// unconstrained scalar (mode 2) with the post-step velocity of U on an n^3 periodic grid.
template <class T> void launch_species_to_fractions(T* s1, T* s2, const T* rho, int64_t nc, T maxt, hipStream_t st);
template <class T>
void launch_backtrace(T* o1, T* o2, const T* a1, const T* a2, Fields<const T> U, int n, T dtdx, T maxt, int mode,
                      hipStream_t st);

template <class T>
void launch_wavespeed(Fields<const T> U, const Grid& G, const StepParams<T>& P, double* red, double* out,
                      hipStream_t st);
template <class T>
void launch_divb(Fields<const T> U, const Grid& G, T inv2dx, double* red, double* out3, hipStream_t st);
template <class T> void launch_totals(Fields<const T> U, const Grid& G, double* red, double* out8, hipStream_t st);
// ct_update! (ct.jl:158-173) on U. With P, red and dsmax: the cell-field pass also folds
// the CFL speed of the finished state into *dsmax (device), as the fused GLM step does.
// have_edges: slots 15-17 already hold this step's edge EMFs (formed by the step kernel).
template <class T>
void launch_ct_update(Fields<T> U, T* ctp, const Grid& G, T dtdx, bool compensate, hipStream_t st,
                      const StepParams<T>* P = nullptr, double* red = nullptr, double* dsmax = nullptr,
                      bool have_edges = false);
template <class T> void launch_divb_face(const T* ctp, const Grid& G, T invdx, double* red, double* out, hipStream_t st);
template <class T>
void launch_ct_install(Fields<T> U, T* ctp, const Grid& G, hipStream_t st);
template <class T>
void launch_ct_curl(const T* Ax, const T* Ay, const T* Az, T* bx, T* by, T* bz, const Grid& G, T invdx,
                    hipStream_t st);
template <class T>
void launch_riemann_batch(const T* L, const T* R, T* F, int64_t n, int dir, bool ct, const StepParams<T>& P,
                          hipStream_t st);

}  // namespace pmf
