// The batched ensemble: many independent periodic float32 GLM cubes of uneven sizes
// advanced together, one launch per stage (per-cube time step, persistent step over
// all cubes' tiles), the step pair captured in a HIP graph, and an autotuner that
// picks the z-chunk length of the tiles per cube size and the persistent grid size.
// The reference advances one MHDState per step! call (driver.jl:20-37); an ensemble
// is equivalent to running step! with evolve!'s dt rule (driver.jl:48-71) on each cube
// separately, which is how it is tested.
#include <algorithm>
#include <cstring>
#include <limits>
#include <map>
#include <vector>

#include "state.hpp"

using namespace pmf;

struct pmf_ensemble {
  pmf_params p{};
  int ncubes = 0;
  std::vector<int> edge, zc;
  float* pool = nullptr;
  std::vector<EnsCube> cubes;
  EnsCube* dcubes = nullptr;
  std::vector<int2> items;
  int2* ditems = nullptr;
  size_t item_cap = 0;
  EnsQueues q{};
  double* red = nullptr;
  StepParams<float>* params = nullptr;
  double* clock = nullptr;  // (t, dt, smax) per cube
  unsigned* ctr = nullptr;
  int nwg = 512;
  int parity = 0;  // 0: the state is in A
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  bool graph_ok = false;
  int g_parity = -1;
  float g_cfl = 0, g_fac = 0, g_cr = 0;
};

namespace {

void drop_graph(pmf_ensemble* e) {
  if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->gexec = nullptr;
  e->graph = nullptr;
  e->graph_ok = false;
}

StepParams<float> base_params(const pmf_ensemble* e) {
  const pmf_params& p = e->p;
  StepParams<float> P{};
  P.gamma = float(p.gamma);
  P.gm1 = float(p.gamma) - 1.f;
  P.igm1 = 1.f / P.gm1;
  P.smallr = float(p.smallr);
  P.pfl = float(p.pfl);
  P.llf_dmin = float(p.llf_dmin);
  P.llf_pmin = float(p.llf_pmin);
  P.decay = 1.f;
  P.recon = p.recon;
  P.rsolve = p.riemann;
  return P;
}

// Item list (cubes largest first, each cube's tiles x, y, then z chunk) and the eight
// queue slices of equal cost (planes marched per item: its chunk length + 2).
void build_items(pmf_ensemble* e) {
  std::vector<int> order(e->ncubes);
  for (int c = 0; c < e->ncubes; ++c) order[c] = c;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return e->edge[a] > e->edge[b]; });
  e->items.clear();
  std::vector<double> cost;
  for (int c : order) {
    EnsCube& cb = e->cubes[c];
    cb.g = ens_geom(e->edge[c], e->zc[c]);
    cb.item0 = (int32_t)e->items.size();
    cb.nitems = cb.g.ntx * cb.g.nty * cb.g.nchunks;
    for (int w = 0; w < cb.nitems; ++w) {
      e->items.push_back(int2{c, w});
      const int chunk = w / (cb.g.ntx * cb.g.nty);
      cost.push_back(double(std::min(cb.g.zc, cb.g.nz - chunk * cb.g.zc) + 2));
    }
  }
  double total = 0;
  for (double x : cost) total += x;
  int start = 0;
  double acc = 0;
  const int n = (int)e->items.size();
  for (int k = 0; k < 8; ++k) {
    int end = start;
    const double goal = total * (k + 1) / 8.0;
    if (k == 7) end = n;
    else
      while (end < n && acc + 0.5 * cost[end] < goal) acc += cost[end++];
    e->q.base[k] = start;
    e->q.len[k] = end - start;
    start = end;
  }
  if (e->items.size() > e->item_cap) {
    (void)hipFree(e->ditems);
    (void)hipFree(e->red);
    e->ditems = nullptr;
    e->red = nullptr;
    e->item_cap = e->items.size();
    hip_check(hipMalloc(&e->ditems, sizeof(int2) * e->item_cap), "hipMalloc");
    hip_check(hipMalloc(&e->red, sizeof(double) * e->item_cap), "hipMalloc");
  }
  // on the ensemble's stream, so the copies are ordered after launches already queued there
  // and before the next ones; the wait keeps the host vectors alive until they are read
  hip_check(hipMemcpyAsync(e->ditems, e->items.data(), sizeof(int2) * e->items.size(), hipMemcpyHostToDevice, e->st),
            "copy");
  hip_check(hipMemcpyAsync(e->dcubes, e->cubes.data(), sizeof(EnsCube) * e->ncubes, hipMemcpyHostToDevice, e->st),
            "copy");
  hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
  drop_graph(e);
}

void check_ens(const pmf_ensemble* e) {
  if (!e) throw Error("null ensemble");
}

const Fields<float>& current(const pmf_ensemble* e, int c) { return e->parity ? e->cubes[c].B : e->cubes[c].A; }

// One ensemble step on the stream: the per-cube CFL stage, then the persistent step.
void enqueue_step(pmf_ensemble* e, int parity, float cfl, float fac, float cr) {
  launch_ens_params(e->dcubes, e->ncubes, e->red, base_params(e), e->params, e->clock, e->ctr, cfl, fac, cr, e->st);
  launch_ens_step(e->dcubes, e->ditems, e->q, e->params, e->red, e->ctr, e->nwg, parity, e->p.recon, e->p.riemann,
                  e->st);
}

void initial_speeds(pmf_ensemble* e) {
  launch_ens_wavespeed(e->dcubes, e->ditems, (int)e->items.size(), base_params(e), e->red, e->parity, e->st);
}

double elapsed_ms(pmf_ensemble* e) {
  hip_check(hipEventSynchronize(e->e1), "hipEventSynchronize");
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, e->e0, e->e1), "hipEventElapsedTime");
  return ms;
}

}  // namespace

extern "C" {

int pmf_ensemble_create(const pmf_params* base, int32_t ncubes, const int32_t* edge, pmf_ensemble** out) {
  return guard([&] {
    if (!base || !edge || !out) throw Error("null argument");
    *out = nullptr;
    if (ncubes <= 0) throw Error("an ensemble needs at least one cube");
    if (base->precision != PMF_F32) throw Error("ensembles run in float32");
    if (base->solver != PMF_SOLVER_GLM) throw Error("ensembles run the GLM solver");
    if (!(base->gamma > 1.0)) throw Error("gamma must exceed 1");
    if (base->riemann < 0 || base->riemann > 2) throw Error("unknown riemann solver code");
    if (base->recon < 0 || base->recon > 2) throw Error("unknown reconstruction code");
    for (int c = 0; c < ncubes; ++c) {
      if (edge[c] < 16 || edge[c] % kFusedTX)
        throw Error("ensemble cube edges must be multiples of 16 (got " + std::to_string(edge[c]) + ")");
      if (edge[c] > 1008)  // each field below 4 GiB: the step addresses cubes with 32-bit offsets
        throw Error("ensemble cube edges must be at most 1008 (got " + std::to_string(edge[c]) + ")");
    }
    auto* e = new pmf_ensemble();
    try {
      e->p = *base;
      e->ncubes = ncubes;
      e->edge.assign(edge, edge + ncubes);
      e->zc.resize(ncubes);
      for (int c = 0; c < ncubes; ++c) e->zc[c] = std::min(16, edge[c]);
      size_t total = 0;
      for (int c = 0; c < ncubes; ++c) total += size_t(18) * edge[c] * edge[c] * edge[c];
      hip_check(hipMalloc(&e->pool, sizeof(float) * total), "hipMalloc");
      e->cubes.resize(ncubes);
      size_t off = 0;
      for (int c = 0; c < ncubes; ++c) {
        const size_t nc = size_t(edge[c]) * edge[c] * edge[c];
        for (int v = 0; v < 9; ++v) e->cubes[c].A.v[v] = e->pool + off + v * nc;
        for (int v = 0; v < 9; ++v) e->cubes[c].B.v[v] = e->pool + off + (9 + v) * nc;
        e->cubes[c].dx = 1.0 / edge[c];
        off += 18 * nc;
      }
      hip_check(hipMemset(e->pool, 0, sizeof(float) * total), "hipMemset");
      hip_check(hipMalloc(&e->dcubes, sizeof(EnsCube) * ncubes), "hipMalloc");
      hip_check(hipMalloc(&e->params, sizeof(StepParams<float>) * ncubes), "hipMalloc");
      hip_check(hipMalloc(&e->clock, sizeof(double) * 3 * ncubes), "hipMalloc");
      hip_check(hipMemset(e->clock, 0, sizeof(double) * 3 * ncubes), "hipMemset");
      hip_check(hipMalloc(&e->ctr, sizeof(unsigned) * 8), "hipMalloc");
      hip_check(hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking), "hipStreamCreate");
      hip_check(hipEventCreate(&e->e0), "hipEventCreate");
      hip_check(hipEventCreate(&e->e1), "hipEventCreate");
      // the fills above ran on the null stream, which the ensemble's stream does not wait for
      hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
      build_items(e);
    } catch (...) {
      pmf_ensemble_destroy(e);
      throw;
    }
    *out = e;
  });
}

int pmf_ensemble_destroy(pmf_ensemble* e) {
  if (!e) return 0;
  drop_graph(e);
  if (e->st) (void)hipStreamSynchronize(e->st);
  (void)hipFree(e->pool);
  (void)hipFree(e->dcubes);
  (void)hipFree(e->ditems);
  (void)hipFree(e->red);
  (void)hipFree(e->params);
  (void)hipFree(e->clock);
  (void)hipFree(e->ctr);
  if (e->e0) (void)hipEventDestroy(e->e0);
  if (e->e1) (void)hipEventDestroy(e->e1);
  if (e->st) (void)hipStreamDestroy(e->st);
  delete e;
  return 0;
}

int pmf_ensemble_upload(pmf_ensemble* e, int32_t cube, int32_t var, const void* host, int64_t n) {
  return guard([&] {
    check_ens(e);
    if (cube < 0 || cube >= e->ncubes || var < 0 || var > 8) throw Error("cube or variable out of range");
    const int64_t nc = int64_t(e->edge[cube]) * e->edge[cube] * e->edge[cube];
    if (n != nc) throw Error("upload size mismatch");
    if (!host) throw Error("null argument");
    hip_check(hipMemcpyAsync(current(e, cube).v[var], host, sizeof(float) * nc, hipMemcpyHostToDevice, e->st),
              "copy");
    hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
  });
}

int pmf_ensemble_download(pmf_ensemble* e, int32_t cube, int32_t var, void* host, int64_t n) {
  return guard([&] {
    check_ens(e);
    if (cube < 0 || cube >= e->ncubes || var < 0 || var > 8) throw Error("cube or variable out of range");
    const int64_t nc = int64_t(e->edge[cube]) * e->edge[cube] * e->edge[cube];
    if (n != nc) throw Error("download size mismatch");
    if (!host) throw Error("null argument");
    hip_check(hipMemcpyAsync(host, current(e, cube).v[var], sizeof(float) * nc, hipMemcpyDeviceToHost, e->st),
              "copy");
    hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
  });
}

int pmf_ensemble_run(pmf_ensemble* e, int32_t nsteps, double cfl, double glm_ch_fac, double glm_cr, int32_t use_graph,
                     double* seconds) {
  return guard([&] {
    check_ens(e);
    if (nsteps < 0) throw Error("nsteps must be non-negative");
    if (!(cfl > 0)) throw Error("cfl must be positive");
    if (!(glm_ch_fac > 0)) throw Error("glm_ch_fac must be positive");
    if (!(glm_cr >= 0)) throw Error("glm_cr must be non-negative");
    const float fcfl = float(cfl), ffac = float(glm_ch_fac), fcr = float(glm_cr);
    initial_speeds(e);
    hip_check(hipEventRecord(e->e0, e->st), "hipEventRecord");
    int k = 0;
    if (use_graph && nsteps >= 2) {
      if (!e->graph_ok || e->g_parity != e->parity || e->g_cfl != fcfl || e->g_fac != ffac || e->g_cr != fcr) {
        drop_graph(e);
        hip_check(hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        enqueue_step(e, e->parity, fcfl, ffac, fcr);
        enqueue_step(e, e->parity ^ 1, fcfl, ffac, fcr);
        hip_check(hipStreamEndCapture(e->st, &e->graph), "hipStreamEndCapture");
        hip_check(hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        e->graph_ok = true;
        e->g_parity = e->parity;
        e->g_cfl = fcfl;
        e->g_fac = ffac;
        e->g_cr = fcr;
      }
      for (; k + 2 <= nsteps; k += 2) hip_check(hipGraphLaunch(e->gexec, e->st), "hipGraphLaunch");
    }
    int par = e->parity;
    for (; k < nsteps; ++k) {
      enqueue_step(e, par, fcfl, ffac, fcr);
      par ^= 1;
    }
    hip_check(hipGetLastError(), "ensemble launch");
    hip_check(hipEventRecord(e->e1, e->st), "hipEventRecord");
    const double ms = elapsed_ms(e);
    e->parity ^= (nsteps & 1);
    if (seconds) *seconds = ms * 1e-3;
  });
}

int pmf_ensemble_clock(pmf_ensemble* e, double* t, double* dt, double* smax) {
  return guard([&] {
    check_ens(e);
    std::vector<double> h(3 * e->ncubes);
    hip_check(hipMemcpyAsync(h.data(), e->clock, sizeof(double) * h.size(), hipMemcpyDeviceToHost, e->st), "copy");
    hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
    for (int c = 0; c < e->ncubes; ++c) {
      if (t) t[c] = h[3 * c];
      if (dt) dt[c] = h[3 * c + 1];
      if (smax) smax[c] = h[3 * c + 2];
    }
  });
}

int pmf_ensemble_config(pmf_ensemble* e, int32_t* zc, int32_t* nwg, int32_t* nitems) {
  return guard([&] {
    check_ens(e);
    if (zc)
      for (int c = 0; c < e->ncubes; ++c) zc[c] = e->zc[c];
    if (nwg) *nwg = e->nwg;
    if (nitems) *nitems = (int32_t)e->items.size();
  });
}

int pmf_ensemble_set_config(pmf_ensemble* e, const int32_t* zc, int32_t nwg) {
  return guard([&] {
    check_ens(e);
    if (nwg < 8 || nwg % 8) throw Error("the persistent grid must be a positive multiple of 8 workgroups");
    if (zc)
      for (int c = 0; c < e->ncubes; ++c)
        if (zc[c] < 1) throw Error("z-chunk lengths must be positive");
    hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
    if (zc)
      for (int c = 0; c < e->ncubes; ++c) e->zc[c] = std::min<int>(zc[c], e->edge[c]);
    e->nwg = nwg;
    build_items(e);
  });
}

// Coordinate search: per distinct cube edge, the z-chunk length in {4, 8, 16, 32, 64}
// (at most the edge), then the persistent grid in {256, 512, 768, 1024} workgroups,
// each candidate timed over `reps` steps that write the inactive field set only (the
// state and the per-cube clocks are left as they were).
int pmf_ensemble_tune(pmf_ensemble* e, int32_t reps, double* ms_per_step) {
  return guard([&] {
    check_ens(e);
    if (reps < 1) throw Error("reps must be positive");
    drop_graph(e);
    std::vector<double> saved(3 * e->ncubes);
    hip_check(hipMemcpyAsync(saved.data(), e->clock, sizeof(double) * saved.size(), hipMemcpyDeviceToHost, e->st),
              "copy");
    hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
    auto measure = [&]() {
      build_items(e);
      initial_speeds(e);
      enqueue_step(e, e->parity, 0.2f, 1.f, 0.18f);  // warm-up; red now holds the output's maxima
      initial_speeds(e);
      hip_check(hipEventRecord(e->e0, e->st), "hipEventRecord");
      for (int r = 0; r < reps; ++r) enqueue_step(e, e->parity, 0.2f, 1.f, 0.18f);
      hip_check(hipEventRecord(e->e1, e->st), "hipEventRecord");
      hip_check(hipGetLastError(), "ensemble launch");
      return elapsed_ms(e) / reps;
    };
    double best = measure();
    std::map<int, std::vector<int>> classes;
    for (int c = 0; c < e->ncubes; ++c) classes[e->edge[c]].push_back(c);
    for (auto& kv : classes) {
      const int n = kv.first;
      int best_zc = e->zc[kv.second[0]];
      for (int cand : {4, 8, 16, 32, 64}) {
        if (cand > n || cand == best_zc) continue;
        for (int c : kv.second) e->zc[c] = cand;
        const double ms = measure();
        if (ms < best) {
          best = ms;
          best_zc = cand;
        }
      }
      for (int c : kv.second) e->zc[c] = best_zc;
    }
    int best_nwg = e->nwg;
    for (int cand : {256, 512, 768, 1024}) {
      if (cand == best_nwg) continue;
      e->nwg = cand;
      const double ms = measure();
      if (ms < best) {
        best = ms;
        best_nwg = cand;
      }
    }
    e->nwg = best_nwg;
    build_items(e);
    hip_check(hipMemcpyAsync(e->clock, saved.data(), sizeof(double) * saved.size(), hipMemcpyHostToDevice, e->st),
              "copy");
    hip_check(hipStreamSynchronize(e->st), "hipStreamSynchronize");
    if (ms_per_step) *ms_per_step = best;
  });
}

}  // extern "C"
