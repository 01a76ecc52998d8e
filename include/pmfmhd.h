/* pmfmhd — GLM-HLLD finite-volume MHD step on MI355X (gfx950), C ABI.
 *
 * A drop-in for the step/flux API of yipihey/pmf-glm lib/MHDKernels: the state
 * container, the per-step update, the CFL reduction, the diagnostics and the
 * face-staggered constrained-transport (CT) variant. Every entry point returns an int
 * status (0 = ok, nonzero = error); pmf_last_error() gives the calling thread's message,
 * which reuses the reference's text where the same misuse is reported there.
 *
 * Layout (state.jl:1-13,95): each field is a flat array of nx*ny*nz values, cell
 * (i,j,k) (0-based) at (k*ny + j)*nx + i. Conserved variables are numbered 0..8 in the
 * reference's order (rho, mx, my, mz, E, Bx, By, Bz, psi). Buffers live on the GPU and
 * are owned by the library; host pointers passed in are plain host memory.
 * Calls on one state are not re-entrant; every call that returns a host value waits
 * for the state's stream, the others are asynchronous on that stream.
 */
#ifndef PMFMHD_H
#define PMFMHD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Codes, numbered as in the reference. */
enum { PMF_RIEMANN_HLLD = 0, PMF_RIEMANN_HLL = 1, PMF_RIEMANN_LLF = 2 };          /* glm_physics.jl:309-311 */
enum { PMF_RECON_PLM = 0, PMF_RECON_PPM = 1, PMF_RECON_PLM_MINMOD = 2 };         /* glm_physics.jl:202-204 */
enum { PMF_BC_PERIODIC = 0, PMF_BC_OUTFLOW = 1, PMF_BC_REFLECTING = 2 };         /* boundary.jl:7-9 */
enum { PMF_SOLVER_GLM = 0, PMF_SOLVER_CT = 1 };                                  /* state.jl:130-131 */
enum { PMF_F32 = 0, PMF_F64 = 1 };
enum { PMF_INTEGRATOR_REF = 0, PMF_INTEGRATOR_CUBE = 1 };                        /* driver.jl:25-35 */

/* Keyword arguments of allocate_state (state.jl:123-129). */
typedef struct pmf_params {
  int32_t nx, ny, nz;
  int32_t riemann;   /* PMF_RIEMANN_* */
  int32_t recon;     /* PMF_RECON_* */
  int32_t bc[3];     /* PMF_BC_* per axis */
  int32_t solver;    /* PMF_SOLVER_* */
  int32_t precision; /* PMF_F32 / PMF_F64: element type of every field */
  double dx, gamma, smallr, pfl, llf_dmin, llf_pmin;
} pmf_params;

typedef struct pmf_state pmf_state;

/* Fills the reference defaults: gamma 5/3, smallr 1e-6, pfl 1e-7, llf_dmin 1e-4,
 * llf_pmin 0, HLLD, PLM (MonCen), periodic, GLM, f32 (state.jl:123-129). */
int pmf_params_default(pmf_params* p, int32_t nx, int32_t ny, int32_t nz, double dx);

/* allocate_state (state.jl:123-141): zero-filled U, scratch and, for CT, the 21-slot
 * face/workspace buffer (state.jl:44-58), on the current HIP device. */
int pmf_create(const pmf_params* p, pmf_state** out);
int pmf_destroy(pmf_state* s);
/* A state without fields, for a caller that owns its arrays (the reference's MHDState
 * holds nine separate device vectors, state.jl:63-65): it carries the parameters, the
 * stream and the reduction workspace used by pmf_step_into / pmf_max_wavespeed_of. */
int pmf_create_workspace(const pmf_params* p, pmf_state** out);

/* Host <-> device copies of one conserved variable (n = number of cells; the host
 * buffer holds elements of the state's precision). fields_to_host: state.jl:144. */
int pmf_upload(pmf_state* s, int32_t var, const void* host, int64_t n);
int pmf_download(pmf_state* s, int32_t var, void* host, int64_t n);
/* Device address of U[var] / scratch[var] for zero-copy wrapping. After pmf_step the
 * scratch set holds the previous state (integrator_ref.jl:217); between steps scratch
 * is caller workspace: the library keeps nothing there. */
int pmf_device_ptr(pmf_state* s, int32_t var, void** ptr);
int pmf_scratch_ptr(pmf_state* s, int32_t var, void** ptr);
/* CT slot q in 0..20 (state.jl:17-37 order: bxf byf bzf, mfx mfy mfz, fxby fxbz fybx
 * fybz fzbx fzby, ecx ecy ecz, ex ey ez, rbx rby rbz). */
int pmf_ct_ptr(pmf_state* s, int32_t slot, void** ptr);
/* Host copy of CT slot q (n = number of cells). */
int pmf_ct_download(pmf_state* s, int32_t slot, void* host, int64_t n);

/* Work on a caller stream (hipStream_t, default priority); NULL restores the state's own. */
int pmf_set_stream(pmf_state* s, void* hip_stream);
int pmf_synchronize(pmf_state* s);

/* max_wavespeed! / compute_dt! (timestep.jl:17-41): smax = max over cells and axes of
 * |v_d| + c_fast,d; dt = cfl*dx/max(smax, eps(T)), computed in the state's precision. */
int pmf_max_wavespeed(pmf_state* s, double* smax);
int pmf_compute_dt(pmf_state* s, double cfl, double* dt, double* smax);

/* step_ref! / step_cube! (integrator_ref.jl:210-220, integrator_cube.jl:455-463): one
 * MUSCL-Hancock step of size dt with GLM speed ch and psi factor decay, written to
 * scratch and swapped in; CT states also run ct_update! (ct.jl:158-173) and ignore ch
 * and decay. integrator: PMF_INTEGRATOR_REF (per-cell recompute) or
 * PMF_INTEGRATOR_CUBE (fused tiled kernel; GLM only; any grid shape and boundary codes,
 * where the reference's needs a cubic N % 4 == 0 grid; a reflecting axis needs two
 * cells). */
int pmf_step_raw(pmf_state* s, double dt, double ch, double decay, int32_t integrator);
/* step! (driver.jl:20-37): decay = exp(-glm_cr*ch*dt/dx) in the state's precision. */
int pmf_step(pmf_state* s, double dt, double ch, double glm_cr, int32_t integrator);
/* _cube_launch! / _launch_step_ref! seam (integrator_cube.jl:467-476,
 * integrator_ref.jl:189-201): one step from the caller's nine input field arrays U[0..8]
 * into its nine output arrays O[0..8] (no swap; the caller swaps as step_ref! does), then
 * for CT states ct_update! on O using the caller's 21-slot packed buffer ct (NULL for
 * GLM). Asynchronous on the state's stream. */
int pmf_step_into(pmf_state* s, const void* const* U, void* const* O, void* ct, double dt, double ch,
                  double decay, int32_t integrator);
/* max_wavespeed! (timestep.jl:17-23) of the caller's nine field arrays. */
int pmf_max_wavespeed_of(pmf_state* s, const void* const* U, double* smax);
/* pmf_step_into without a state and without allocating. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, hosts that forbid hipMalloc). p gives the grid and the
 * physics as for pmf_create; the step runs on hip_stream (NULL = the null stream) and its
 * kernels' workspace is a library-held device array of 2^18 + 64 doubles (2 MiB), one per
 * device, kept until the process ends: grids whose fused step needs more than 2^18
 * per-workgroup partial slots (about 1.6e9 cells at the built-in chunk) are refused with an
 * error naming pmf_step_into. smax_dev (optional, a caller device double) receives the
 * CFL speed of the output fields (max_wavespeed!, timestep.jl:17-23). Calls that overlap
 * on one device share the workspace and race; calls on one stream are ordered. */
int pmf_step_into_static(const pmf_params* p, const void* const* U, void* const* O, void* ct, double dt, double ch,
                         double decay, int32_t integrator, void* hip_stream, double* smax_dev);
/* evolve! (driver.jl:48-71) without a callback: CFL steps until t >= tfinal(1-1e-9),
 * ch = glm_ch_fac*smax, dt scaled by smax/max(smax,ch), the last step clipped. */
int pmf_evolve(pmf_state* s, double tfinal, double cfl, double glm_ch_fac, double glm_cr,
               int32_t integrator, double* t_out, int64_t* nsteps_out);

/* nsteps steps of evolve!'s rule (driver.jl:48-71: dt = cfl*dx/smax, ch =
 * glm_ch_fac*smax, dt scaled by smax/max(smax,ch), psi decay exp(-glm_cr*ch*dt/dx)) on the
 * fused kernel, with each step's dt computed on the device from the speed the previous
 * step folded: the host enqueues every step without waiting (GLM states). Each step is
 * the pmf_compute_dt + pmf_step pair bit for bit. t_out (optional) receives the elapsed
 * time and dt_out the last step's dt; the call waits for the device once, at the end.
 * No counterpart in the reference, whose driver computes every dt on the host. */
int pmf_advance(pmf_state* s, int64_t nsteps, double cfl, double glm_ch_fac, double glm_cr, double* t_out,
                double* dt_out);

/* Diagnostics. max_divb (diagnostics.jl:6-36): BC-aware centred |div B| maximum.
 * pmf_divb_norm_fp64: the same divergence's float64 L2 norm sqrt(mean(d^2)) and max,
 * with compensated summation. conserved_totals (diagnostics.jl:44-55): float64
 * compensated sums x dV of rho, mx, my, mz, E, Bx, By, Bz (psi excluded). */
int pmf_max_divb(pmf_state* s, double* out);
int pmf_divb_norm_fp64(pmf_state* s, double* l2, double* linf);
int pmf_conserved_totals_fp64(pmf_state* s, double out[8]);

/* CT: max_divb_face (ct.jl:176-201); _ct_install_faces! (ct.jl:228-236: faces in,
 * remainders zeroed, cell B = face average, E keeps the thermal pressure, psi = 0);
 * ct_faces_from_potential! curl step (ct.jl:310-320) followed by the install. Host
 * arrays of nx*ny*nz elements in the state's precision. */
int pmf_max_divb_face(pmf_state* s, double* out);
int pmf_ct_install_faces(pmf_state* s, const void* bxf, const void* byf, const void* bzf);
int pmf_ct_faces_from_potential(pmf_state* s, const void* ax, const void* ay, const void* az);

/* riemann / riemann_ct (glm_physics.jl:317-348) for n faces: L, R, F are host arrays of
 * n x 9 primitive states / fluxes (row-major, state precision); dir 0,1,2 = x,y,z. */
int pmf_riemann_flux_batch(pmf_state* s, const void* L, const void* R, int32_t dir, double ch,
                           int64_t n, void* F);

/* Tile depth of the fused (:cube) kernel: planes per z chunk, at least 4 (0 = the
 * built-in choice, the default). Results do not depend on it. */
int pmf_set_chunk(pmf_state* s, int32_t zc);
/* Times the fused kernel (the plain step, or the fused CT step for CT states) for z
 * chunks of 4..128 planes (reps launches each, writing only the scratch fields and, for
 * CT, the per-step workspace slots) and keeps the fastest; returns it and its time per
 * step. */
int pmf_autotune(pmf_state* s, int32_t reps, int32_t* zc_out, double* ms_out);
/* Kernel of the CT (:ref) step. By default a constrained-transport state steps (plain,
 * drag and radiation steps alike) on the fused z-marching kernel, which solves every face
 * once with the :ref algorithm (integrator_ref.jl:83-139, 222-258); per_cell != 0
 * selects the per-cell kernel of integrator_ref.jl:189-220 instead. No counterpart in the
 * reference (one kernel there); results agree to rounding. */
int pmf_set_ct_kernel(pmf_state* s, int32_t per_cell);
/* Workspace slots of the fused CT step. The reference's step writes every face's mass and
 * transverse-field fluxes and every cell's EMF to slots 3-14 of the CTState buffer, and
 * ct_update! forms the edge EMFs (slots 15-17) from them (integrator_ref.jl:83-139,
 * ct.jl:71-93). By default (fill = 0) the fused float PLM / MonCen step forms the edge EMFs
 * itself, bit for bit as that pass would, and leaves slots 3-14 as they were: every
 * reader of those slots in the reference writes them first. fill != 0 writes slots 3-14
 * and runs the separate edge pass, as the reference does. No counterpart in the reference
 * (its step always writes them); cells, faces, edge EMFs and remainders are the same
 * either way. */
int pmf_set_ct_workspace(pmf_state* s, int32_t fill);
/* Kernel timing: while enabled, HIP events bracket every step-kernel launch on the
 * state's stream; pmf_timing_read waits for the stream and returns the summed kernel
 * time (ms) and the number of timed launches since the last enable. */
int pmf_timing_enable(pmf_state* s, int32_t on);
int pmf_timing_read(pmf_state* s, double* total_ms, int64_t* launches);

/* ---- Learned face flux (a BF16 MLP on MFMA standing in for HLLD on smooth faces) ----
 * No counterpart in the reference: it adds a learned flux on top of the MHDKernels
 * step. Network 16 -> 64 -> 64 -> 64 -> 9 (tanh GELU), float32 master weights, bfloat16
 * operands with float32 accumulation. Inputs: the face's mean and jump of the 8
 * normal-frame primitives, normalised by the local fast-speed scale (jumps divided by the
 * gate threshold); outputs: the HLLD dissipation relative to the central flux. The gate
 * (jump detector) admits a face when the relative jumps of rho, p and |B| are below thr. */
typedef struct pmf_surrogate pmf_surrogate;
int pmf_surrogate_create(pmf_state* s, uint64_t seed, pmf_surrogate** out);  /* s: float32; stream + params */
int pmf_surrogate_destroy(pmf_surrogate* n);
int pmf_surrogate_nparams(void);                                              /* 9993 */
int pmf_surrogate_get_weights(pmf_surrogate* n, float* host, int64_t count);  /* [W1 b1 W2 b2 W3 b3 W4 b4] */
int pmf_surrogate_set_weights(pmf_surrogate* n, const float* host, int64_t count);
/* Training set on the device: capacity, dump from a state (every gated lower face of
 * every cell kept with probability keep; dt, ch as for the step), copy out / upload.
 * Rows: x[16], y[9], state[18] = normal-frame L[8], R[8], GLM flux of Bn, of psi. */
int pmf_surrogate_reserve(pmf_surrogate* n, int64_t capacity);
int pmf_surrogate_dump(pmf_surrogate* n, pmf_state* s, double dt, double ch, double thr, double keep,
                       uint64_t seed, int64_t* total);
int pmf_surrogate_dataset(pmf_surrogate* n, int64_t first, int64_t count, float* x, float* y, float* state);
int pmf_surrogate_upload(pmf_surrogate* n, const float* x, const float* y, const float* state, int64_t count);
/* Full-batch loss gradient over a dataset range (bf16 = 1: bfloat16 MFMA, 0: float32 MFMA). */
int pmf_surrogate_grad(pmf_surrogate* n, int64_t first, int64_t count, int32_t bf16, float* grad, double* loss);
int pmf_surrogate_predict(pmf_surrogate* n, int64_t first, int64_t count, int32_t bf16, float* y);
/* Adam on hashed mini-batches; losses (optional) receives the loss of every step. */
int pmf_surrogate_train(pmf_surrogate* n, int32_t batch, int32_t steps, double lr, uint64_t seed, int32_t bf16,
                        double* losses);
/* Face fluxes of dataset rows: mode 0 HLLD, 1 surrogate, 2 central flux only. */
int pmf_surrogate_face_flux(pmf_surrogate* n, int64_t first, int64_t count, int32_t mode, double thr, float* F);
/* step! with the fused kernel in which every face whose jumps pass the gate (thr) takes
 * the learned flux and every other face HLLD (float32, GLM, HLLD, periodic, nx % 16 == 0).
 * nsmooth (optional; waits for the stream) receives the number of network-solved faces. */
int pmf_step_surrogate(pmf_state* s, pmf_surrogate* n, double dt, double ch, double glm_cr, double thr,
                       int64_t* nsmooth);

/* ---- Drag / radiation-coupled steps -----------------------------------------------
 * step_ref_drag! / step_cube_drag! (integrator_ref.jl:222-236, integrator_cube.jl:478-503):
 * the step with the Compton-drag propagator of impulse q = Gamma dt toward the target
 * velocity (NULL = rest; the :cube integrator takes rest only). Swaps like pmf_step. */
int pmf_step_drag(pmf_state* s, double dt, double ch, double decay, double drag_impulse,
                  const double* target_velocity, int32_t integrator);
/* step_ref_radiation! / step_cube_radiation! (integrator_ref.jl:238-258,
 * integrator_cube.jl:505-540): drag plus the radiation correction buffer, a device array of
 * 9 slots of ncells in the state's precision: psi (read instead of the state's psi), the
 * three velocity corrections added to every face state of a cell, then the density
 * perturbation (or fallback mask), three field and one energy residual updated in place
 * when tracked. flags: PMF_SRC_*; the check and fallback options are :cube only. */
enum {
  PMF_SRC_TRACK_DENSITY = 1,
  PMF_SRC_TRACK_MAGNETIC = 2,
  PMF_SRC_TRACK_ENERGY = 4,
  PMF_SRC_CHECK_ADMISSIBILITY = 8, /* inadmissible cells come back with a negative density */
  PMF_SRC_FALLBACK = 16,           /* cells with slot 4 > 0: zero slopes, LLF on their faces */
  PMF_SRC_GLOBAL_FALLBACK = 32,    /* every cell first order, LLF everywhere */
  PMF_SRC_ROBUST_FALLBACK = 64     /* minmod PLM and HLL for this step */
};
int pmf_step_radiation(pmf_state* s, double dt, double ch, double decay, double drag_impulse, void* correction,
                       int32_t flags, int32_t integrator);
/* The _cube_drag_launch! / _cube_radiation_launch! seam: one coupled step from nine
 * caller arrays into nine others (no swap); radiation when correction != NULL. */
int pmf_step_sources_into(pmf_state* s, const void* const* U, void* const* O, void* ct, void* correction, double dt,
                          double ch, double decay, double drag_impulse, const double* target_velocity,
                          int32_t flags, int32_t integrator);

/* ---- State handoff files ----------------------------------------------------------
 * The production runner's raw handoff (src/production/runner.jl:3814-3883): in dir, the
 * files mhd_U_1.f32 ... mhd_U_9.f32 hold the nine conserved variables as nx*ny*nz raw
 * little-endian float32 values each, x fastest. Import: every file must hold exactly
 * that many bytes ("bad raw Float32 size for <path>: got <n> bytes, expected <m>"; a
 * missing file reads as 0 bytes) and the minimum density and energy must be positive
 * and finite ("imported MHD state has invalid minimum density <x>" / "... energy <x>");
 * all checks run before the state is touched. float64 states convert on the way in and
 * out. Export: each file written through <path>.tmp and a rename. */
int pmf_import_mhd_raw(pmf_state* s, const char* dir);
int pmf_export_mhd_raw(pmf_state* s, const char* dir);

/* ---- Passive species (passive_transport.jl) --------------------------------------
 * conserved_species_to_fraction(s)!: species densities to bounded mass fractions in
 * place, using rho (NULL: the state's density); with two species the second bounds the
 * first (q1 <= maximum_total - q2). species2 may be NULL. */
int pmf_species_to_fractions(pmf_state* s, void* species1, void* species2, const void* rho, double maximum_total);
/* advect_species_fraction(s)!: backtrace the fractions with the state's post-step cell
 * velocity (cubic periodic grid, trilinear), restore mass densities with the new gas
 * density, via the caller's scratch fields. species2 / scratch2 may be NULL. */
int pmf_advect_species(pmf_state* s, void* species1, void* species2, void* scratch1, void* scratch2, double dt,
                       double maximum_total);
/* advect_passive_scalar!: the same backtrace of an unconstrained scalar (no bounds, no density). */
int pmf_advect_passive_scalar(pmf_state* s, void* scalar, void* scratch, double dt);

/* Device buffers for callers without a GPU array library (e.g. the radiation
 * correction buffer): zero-filled allocation, free, and copies (kind 1 host to device,
 * 2 device to host, 3 device to device; the copies wait for the device). */
int pmf_buffer_alloc(int64_t bytes, void** out);
int pmf_buffer_free(void* p);
int pmf_buffer_copy(void* dst, const void* src, int64_t bytes, int32_t kind);

/* ---- Ensembles ------------------------------------------------------------------
 * No counterpart in the reference, which advances one MHDState per step! call
 * (driver.jl:20-37): many independent periodic float32 GLM cubes of uneven sizes
 * (edges multiples of 16, dx = 1/edge) advanced together, one launch per stage (the
 * per-cube CFL step of evolve!, driver.jl:48-71, computed on the device; a persistent
 * step over all cubes' tiles), the step pair captured in a HIP graph. base supplies
 * gamma, the floors, llf_dmin/llf_pmin, recon and riemann (nx, ny, nz, dx, bc ignored). */
typedef struct pmf_ensemble pmf_ensemble;
int pmf_ensemble_create(const pmf_params* base, int32_t ncubes, const int32_t* edge, pmf_ensemble** out);
int pmf_ensemble_destroy(pmf_ensemble* e);
/* Conserved variable var of the cube's current state (edge^3 float32 values, i fastest). */
int pmf_ensemble_upload(pmf_ensemble* e, int32_t cube, int32_t var, const void* host, int64_t n);
int pmf_ensemble_download(pmf_ensemble* e, int32_t cube, int32_t var, void* host, int64_t n);
/* nsteps steps of every cube, each with its own dt = cfl dx / smax, ch = glm_ch_fac smax,
 * dt_used = dt smax / max(smax, ch) and psi decay exp(-glm_cr ch dt_used / dx). use_graph
 * replays a captured two-step graph. seconds (optional) = GPU time of the steps. */
int pmf_ensemble_run(pmf_ensemble* e, int32_t nsteps, double cfl, double glm_ch_fac, double glm_cr, int32_t use_graph,
                     double* seconds);
/* Per-cube elapsed time, last dt and the CFL speed the last dt came from (arrays of ncubes). */
int pmf_ensemble_clock(pmf_ensemble* e, double* t, double* dt, double* smax);
/* Autotuner: chooses the tile z-chunk length per cube edge and the persistent grid size
 * by timing reps steps per candidate (the state is not advanced). */
int pmf_ensemble_tune(pmf_ensemble* e, int32_t reps, double* ms_per_step);
int pmf_ensemble_config(pmf_ensemble* e, int32_t* zc, int32_t* nwg, int32_t* nitems);
int pmf_ensemble_set_config(pmf_ensemble* e, const int32_t* zc, int32_t nwg);

const char* pmf_last_error(void);
const char* pmf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* PMFMHD_H */
