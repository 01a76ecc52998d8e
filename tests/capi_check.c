/* A C99 consumer of the boundary: includes only include/pmfmhd.h and links
 * libpmfmhd.so, as a Julia ccall / cgo / JNI binding would. Without arguments it checks
 * what needs no GPU (the defaults, error codes and messages); with "gpu" it also runs a
 * uniform periodic state through compute_dt and one step on both integrators, where
 * every face flux cancels, so the state must come back bit for bit. */
#include <stdio.h>
#include <string.h>

#include "pmfmhd.h"

#define N 16

static int fail(const char* what) {
  fprintf(stderr, "capi_check: %s (%s)\n", what, pmf_last_error());
  return 1;
}

int main(int argc, char** argv) {
  const int gpu = argc > 1 && strcmp(argv[1], "gpu") == 0;
  pmf_params p, bad;
  pmf_state* s = NULL;
  printf("%s\n", pmf_version());
  if (pmf_params_default(&p, N, N, N, 1.0 / N) != 0) return fail("params_default");
  if (p.riemann != PMF_RIEMANN_HLLD || p.recon != PMF_RECON_PLM || p.solver != PMF_SOLVER_GLM) return fail("defaults");
  bad = p;
  bad.nx = 0;
  if (pmf_create(&bad, &s) == 0) return fail("a zero-sized grid was accepted");
  if (!strstr(pmf_last_error(), "grid dimensions must be positive")) return fail("message");
  bad = p;
  bad.solver = PMF_SOLVER_CT;
  bad.bc[0] = PMF_BC_OUTFLOW;
  if (pmf_create(&bad, &s) == 0) return fail("CT with an outflow boundary was accepted");
  if (!strstr(pmf_last_error(), "periodic")) return fail("message");
  if (!gpu) {
    printf("capi_check: ok (no GPU calls)\n");
    return 0;
  }
  {
    static float U[9][N * N * N], V[N * N * N];
    const float rho = 1.3f, vx = 0.4f, vy = -0.2f, vz = 0.3f, pr = 0.8f, bx = 0.5f, by = 0.2f, bz = -0.1f;
    const float e = pr / (5.0f / 3.0f - 1.0f) + 0.5f * rho * (vx * vx + vy * vy + vz * vz) +
                    0.5f * (bx * bx + by * by + bz * bz);
    const float vals[9] = {rho, rho * vx, rho * vy, rho * vz, e, bx, by, bz, 0.0f};
    int integ, v, c, errors = 0;
    for (v = 0; v < 9; ++v)
      for (c = 0; c < N * N * N; ++c) U[v][c] = vals[v];
    for (integ = PMF_INTEGRATOR_REF; integ <= PMF_INTEGRATOR_CUBE; ++integ) {
      double dt, smax;
      if (pmf_create(&p, &s) != 0) return fail("create");
      for (v = 0; v < 9; ++v)
        if (pmf_upload(s, v, U[v], N * N * N) != 0) return fail("upload");
      if (pmf_compute_dt(s, 0.3, &dt, &smax) != 0 || !(dt > 0.0) || !(smax > 0.0)) return fail("compute_dt");
      if (pmf_step(s, dt, smax, 0.18, integ) != 0) return fail("step");
      for (v = 0; v < 9; ++v) {
        double worst = 0.0;
        int nbad = 0;
        if (pmf_download(s, v, V, N * N * N) != 0) return fail("download");
        for (c = 0; c < N * N * N; ++c)
          if (V[c] != U[v][c]) {
            const double d = V[c] > U[v][c] ? (double)V[c] - U[v][c] : (double)U[v][c] - V[c];
            worst = d > worst ? d : worst;
            ++nbad;
          }
        if (nbad) {
          fprintf(stderr, "capi_check: integrator %d, variable %d: %d cells changed, max |change| %.3g\n", integ, v,
                  nbad, worst);
          errors += nbad;
        }
      }
      if (pmf_step(s, dt, 0.0, 0.18, integ) == 0) return fail("ch = 0 was accepted for a GLM state");
      if (!strstr(pmf_last_error(), "needs the GLM cleaning speed")) return fail("message");
      if (pmf_destroy(s) != 0) return fail("destroy");
    }
    if (errors) return 1;
  }
  printf("capi_check: ok (uniform state unchanged by one step on both integrators)\n");
  return 0;
}
