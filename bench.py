"""Throughput of the plain GLM-HLLD step on one MI355X (the JSON line's value), with the
float32-vs-float64 error report, the surrogate-assisted step against the plain step,
and the batched ensemble as extra objects of the same line.

Workload: the 256^3 float32 benchmark configuration (PLM minmod, HLLD, CFL 0.2,
gamma 5/3, periodic unit box; seeded solenoidal Gaussian B with P(k)~k^2, |m|<=32,
rms|B|=1, and Mach-0.5 solenoidal velocity with P(k)~k^-2). One step = the CFL speed
of the current state (fused into the previous step's epilogue on the fused path),
dt = cfl*dx/smax, ch = smax, and one GLM-HLLD update of all cells, with the state
resident in HBM. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "pmf-glm_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "cell-updates/s at 256³ FP32: surrogate vs plain HLLD at rel-L2≤1e-4 vs FP64"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = 72  # 9 float32 read + 9 float32 written per cell-update
# VALU issue capacity without packed FP32: 256 CUs x 64 lanes (four 16-lane SIMDs) per
# cycle at the 2.4 GHz peak engine clock
VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(U, n, steps):
    """The float32 CPU oracle (OpenMP over the host cores) on the same initial state."""
    import numpy as np

    import oracle as O

    s = O.OracleState((n, n, n), 1.0 / n, dtype=np.float32, recon="plm_minmod")
    s.set_conserved(U)
    dt, smax = s.compute_dt(0.2)
    t0 = time.perf_counter()
    for _ in range(steps):
        dt, smax = s.compute_dt(0.2)
        s.step(float(dt), ch=smax)
    el = time.perf_counter() - t0
    return dict(value=n ** 3 * steps / el, unit="cell-updates/s", cores=O.num_threads(), kind="port",
                sample=f"{steps} step(s) of the {n}^3 float32 workload (CFL reduction + per-cell step), "
                       f"C++ restatement of the reference :ref kernel, {el:.1f} s")


VARS = ("rho", "mx", "my", "mz", "E", "bx", "by", "bz")


def rel_l2(U, Uref):
    """Per-variable relative L2 of U against Uref (9, nc), and psi in absolute terms."""
    import numpy as np
    out = {}
    for v, name in enumerate(VARS):
        d = np.linalg.norm(U[v].astype(np.float64) - Uref[v])
        out[name] = float(d / max(np.linalg.norm(Uref[v]), 1e-300))
    num = np.linalg.norm((U[:8].astype(np.float64) - Uref[:8]).ravel())
    out["all_but_psi"] = float(num / np.linalg.norm(Uref[:8].ravel()))
    out["psi_rms_abs"] = float(np.sqrt(np.mean((U[8].astype(np.float64) - Uref[8]) ** 2)))
    return out


def fp64_sequence(M, U, n, steps, cfl=0.2):
    """Run the float64 GPU path (equal to the float64 CPU oracle to 1e-12 in the parity
    tests) for `steps` CFL steps; returns the final state and the (dt, ch) sequence."""
    s = M.allocate_state((n, n, n), 1.0 / n, dtype="float64", recon="plm_minmod")
    s.set_conserved(U)
    seq = []
    for _ in range(steps):
        dt, sm = M.compute_dt(s, cfl)
        seq.append((float(dt), float(sm)))
        M.step(s, float(dt), ch=float(sm), integrator="cube")
    out = s.conserved()
    del s
    return out, seq


def replay(M, U, n, seq, surrogate=None, thr=None):
    """Float32 replay of a (dt, ch) sequence, plain or surrogate-assisted; returns the
    final state, wall seconds of the step loop and the network-solved face counts of the
    first and last step."""
    import numpy as np
    from pmfmhd.surrogate import Surrogate, step_surrogate
    s = M.allocate_state((n, n, n), 1.0 / n, dtype=np.float32, recon="plm_minmod")
    s.set_conserved(U)
    net = None
    if surrogate is not None:
        net = Surrogate(s, seed=0)
        net.set_weights(surrogate)
    counts = []
    s.synchronize()
    t0 = time.perf_counter()
    for k, (dt, ch) in enumerate(seq):
        if net is None:
            M.step(s, dt, ch=ch, integrator="cube")
        else:
            c = step_surrogate(s, net, dt, ch, glm_cr=M.GLM_CR, thr=thr, count=(k == 0 or k == len(seq) - 1))
            if c is not None:
                counts.append(c)
    s.synchronize()
    el = time.perf_counter() - t0
    return s.conserved(), el, counts


def error_report(M, P, n, steps):
    """Float32 plain step against float64 on the same (dt, ch) sequence (oracle-dt replay)."""
    U = P.config_state("turb_256", seed=1, n=n)
    ref, seq = fp64_sequence(M, U, n, steps)
    got, _, _ = replay(M, U, n, seq)
    return {"workload": f"plain GLM-HLLD {n}^3, seed 1, {steps} CFL-0.2 steps", "steps": steps,
            "reference": "float64 GPU fused kernel (equal to the float64 CPU oracle to 1e-12 in the parity tests and "
                         "in fp64_oracle_check), its dt sequence replayed in float32",
            "rel_l2_fp32_vs_fp64": rel_l2(got, ref)}


def oracle_error_report(M, P, n, steps):
    """The float32 GPU run on the benchmark recipe at n^3 (its own CFL dt sequence, the
    fused step) against the float64 CPU oracle replaying that (dt, ch) sequence from the
    same initial state: the whole-run relative L2 per variable. The oracle is the C++
    restatement of the reference's :ref kernel, OpenMP over the host cores."""
    import numpy as np
    import oracle as O
    U = P.config_state("turb_256", seed=1, n=n)
    s = M.allocate_state((n, n, n), 1.0 / n, dtype=np.float32, recon="plm_minmod")
    s.set_conserved(U)
    seq = []
    for _ in range(steps):
        dt, sm = M.compute_dt(s, 0.2)
        seq.append((float(dt), float(sm)))
        M.step(s, float(dt), ch=float(sm), integrator="cube")
    got = s.conserved()
    del s
    o = O.OracleState((n, n, n), 1.0 / n, dtype=np.float64, recon="plm_minmod")
    o.set_conserved(U)
    t0 = time.perf_counter()
    for k, (dt, ch) in enumerate(seq):
        o.step(dt, ch=ch)
        if (k + 1) % 50 == 0:
            log(f"bench: float64 CPU oracle replay, step {k + 1}/{steps} ({time.perf_counter() - t0:.0f} s)")
    el = time.perf_counter() - t0
    return {"workload": f"plain GLM-HLLD {n}^3 (the benchmark recipe with |m| <= N/8), seed 1, {steps} CFL-0.2 "
                        f"steps",
            "steps": steps,
            "reference": "float64 CPU oracle (C++ restatement of the reference :ref kernel), replaying the float32 "
                         "GPU run's (dt, ch) sequence",
            "rel_l2_fp32_vs_fp64": rel_l2(got, o.U), "oracle_seconds": el, "oracle_cores": O.num_threads()}


def fp64_oracle_check(M, P, n, steps):
    """The float64 GPU step against the float64 CPU oracle at full size: the GPU run's
    (dt, ch) sequence replayed by the oracle (OpenMP over the host cores)."""
    import numpy as np
    import oracle as O
    U = P.config_state("turb_256", seed=1, n=n)
    gpu, seq = fp64_sequence(M, U, n, steps)
    s = O.OracleState((n, n, n), 1.0 / n, dtype=np.float64, recon="plm_minmod")
    s.set_conserved(U)
    t0 = time.perf_counter()
    for dt, ch in seq:
        s.step(dt, ch=ch)
    el = time.perf_counter() - t0
    diff = np.abs(gpu - s.U).max(axis=1) / np.maximum(np.abs(s.U).max(axis=1), 1e-300)
    return {"workload": f"plain GLM-HLLD float64 {n}^3, seed 1, {steps} steps, oracle replaying the GPU dt sequence",
            "max_rel_diff_per_variable": [float(x) for x in diff], "oracle_seconds": el,
            "oracle_cores": O.num_threads()}


def surrogate_leg(M, P, n, steps, weights, thr):
    """The surrogate-assisted step against the plain step on the 32^3-recipe inputs at n^3
    (rms |B| 0.3, |m| <= 8, v = 0): throughput of both, and each one's whole-run error
    against float64 on the same (dt, ch) sequence."""
    import numpy as np
    U = P.config_state("bstatic_256", seed=1, n=n)
    ref, seq = fp64_sequence(M, U, n, steps)
    plain, t_plain, _ = replay(M, U, n, seq)
    sur, t_sur, counts = replay(M, U, n, seq, surrogate=weights, thr=thr)
    cells = n ** 3
    e_plain, e_sur = rel_l2(plain, ref), rel_l2(sur, ref)
    # the same run at tighter gates: fewer network faces, smaller whole-run error
    sweep = []
    for t in (thr / 2, thr / 4, thr / 8):
        st, tt, ct = replay(M, U, n, seq, surrogate=weights, thr=t)
        et = rel_l2(st, ref)["all_but_psi"]
        sweep.append({"thr": t, "value": cells * steps / tt, "network_face_fraction_first_last":
                      [c / (3 * cells) for c in ct], "rel_l2_all_but_psi": et, "meets_rel_l2_1e-4": bool(et <= 1e-4)})
        del st
    return {"workload": f"surrogate-assisted GLM-HLLD FP32 {n}^3 on the 32^3-recipe inputs (|m|<=8, rms|B| 0.3, "
                        f"v=0), seed 1, {steps} steps, gate thr {thr}",
            "value": cells * steps / t_sur, "plain_value": cells * steps / t_plain,
            "speedup_vs_plain": t_plain / t_sur,
            "network_face_fraction_first_last": [c / (3 * cells) for c in counts],
            "rel_l2_vs_fp64": {"plain": e_plain, "surrogate": e_sur},
            "meets_rel_l2_1e-4": bool(e_sur["all_but_psi"] <= 1e-4),
            "gate_sweep": sweep,
            "fastest_meeting_rel_l2_1e-4": fastest_meeting(cells * steps / t_sur, e_sur["all_but_psi"], thr, sweep,
                                                           cells * steps / t_plain)}


def surrogate_profile(M, P, n, steps, warmup, weights, thr):
    """Surrogate-assisted steps only (the profiled kernel), on the surrogate workload's
    inputs, each step with its own CFL dt; the network-solved face fraction of the last."""
    import numpy as np
    from pmfmhd.surrogate import Surrogate, step_surrogate
    s = M.allocate_state((n, n, n), 1.0 / n, dtype=np.float32, recon="plm_minmod")
    s.set_conserved(P.config_state("bstatic_256", seed=1, n=n))
    net = Surrogate(s, seed=0)
    net.set_weights(np.load(weights))
    for _ in range(warmup):
        dt, sm = M.compute_dt(s, 0.2)
        step_surrogate(s, net, float(dt), float(sm), thr=thr)
    s.synchronize()
    t0 = time.perf_counter()
    nsm = 0
    for k in range(steps):
        dt, sm = M.compute_dt(s, 0.2)
        nsm = step_surrogate(s, net, float(dt), float(sm), thr=thr, count=k == steps - 1)
    s.synchronize()
    el = time.perf_counter() - t0
    return {"workload": f"surrogate-assisted step only, {n}^3 surrogate inputs, gate {thr}", "steps": steps,
            "ms_per_step": el / max(steps, 1) * 1e3, "network_face_fraction_last": (nsm or 0) / (3 * n ** 3)}


def fastest_meeting(value, err, thr, sweep, plain_value):
    """The fastest gate whose whole-run error meets the 1e-4 target, against plain."""
    rows = [{"thr": thr, "value": value, "rel_l2_all_but_psi": err}] + \
        [{"thr": r["thr"], "value": r["value"], "rel_l2_all_but_psi": r["rel_l2_all_but_psi"]} for r in sweep]
    ok = [r for r in rows if r["rel_l2_all_but_psi"] <= 1e-4]
    if not ok:
        return None
    best = max(ok, key=lambda r: r["value"])
    return dict(best, plain_value=plain_value, speedup_vs_plain=best["value"] / plain_value)


def ensemble_leg(P):
    """The 64-cube ensemble (32 x 32^3, 24 x 64^3, 8 x 128^3), autotuned, 100 steps."""
    import numpy as np
    from pmfmhd.ensemble import Ensemble
    mem = P.ensemble_members()
    e = Ensemble([m["n"] for m in mem])
    for c, m in enumerate(mem):
        e.upload(c, P.ensemble_state(m))
    ms = e.tune(reps=3)
    zc, nwg, nitems = e.config()
    sec = e.run(100)
    cells = sum(m["n"] ** 3 for m in mem)
    return {"workload": "ensemble: 32 x 32^3 + 24 x 64^3 + 8 x 128^3 FP32 GLM-HLLD, own CFL dt per cube, 100 steps, "
                        "one persistent launch per stage, two-step HIP graph",
            "cubes": len(mem), "cells": cells, "value": cells * 100 / sec, "unit": "cell-updates/s",
            "ms_per_step": sec * 10, "tuned_ms_per_step": ms,
            "tuned": {"zc_by_edge": {str(m["n"]): int(z) for m, z in zip(mem, zc)}, "workgroups": nwg,
                      "work_items": nitems}}


def timed_run(M, s, steps, integ, cfl=0.2):
    """`steps` CFL steps (evolve!'s rule with glm_ch_fac 1); wall seconds."""
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dt, sm = M.compute_dt(s, cfl)
        if s.uses_ct:
            M.step(s, float(dt), integrator=integ)
        else:
            M.step(s, float(dt), ch=float(sm), glm_cr=M.GLM_CR, integrator=integ)
    s.synchronize()
    return time.perf_counter() - t0


def ledger(M, s):
    tot = M.conserved_totals(s)
    l2, linf = M.divb_norm(s)
    return {"totals": tot, "divb_l2": l2, "divb_linf": linf}


def mach15_leg(M, P, steps):
    """512^3 plain step (Mach-1.5 turbulence, lognormal density): throughput and the
    float64 compensated ledger and div B norm before and after."""
    import numpy as np
    n = 512
    U = P.config_state("mach15_512", seed=1)
    s = M.allocate_state((n, n, n), 1.0 / n, dtype=np.float32, recon="plm_minmod")
    s.set_conserved(U)
    del U
    before = ledger(M, s)
    timed_run(M, s, 2, "cube")  # warm-up (compiles nothing; first-touch of the scratch set)
    s.timing(True)
    el = timed_run(M, s, steps, "cube")
    kms, nl = s.timing_read()
    s.timing(False)
    after = ledger(M, s)
    drift = {k: (after["totals"][k] - before["totals"][k]) / max(abs(before["totals"][k]), 1e-300)
             for k in ("mass", "energy")}
    return {"workload": f"plain GLM-HLLD FP32 512^3 (Mach 1.5 solenoidal v, lognormal rho sigma 0.3, |m|<=64, "
                        f"rms|B| 1), {steps + 2} steps", "value": n ** 3 * steps / el,
            "kernel_ms": kms / max(nl, 1), "achieved_GBs": 72 * n ** 3 / (kms / max(nl, 1) * 1e-3) / 1e9,
            "ledger_before": before, "ledger_after": after, "relative_drift": drift,
            "finite": bool(all(np.isfinite(a).all() for a in s.fields_to_host()))}


def ct_leg(M, P, n, steps):
    """The CT variant (fused z-marching step kernel + GS05 EMF + face curl) on the same
    inputs as the plain run, after the divergence-free face projection; the per-cell
    kernel (the reference's structure) is timed on the same state for comparison."""
    import numpy as np
    U = P.config_state("turb_256", seed=1, n=n)
    s = M.allocate_state((n, n, n), 1.0 / n, dtype=np.float32, recon="plm_minmod", mhd_solver="ct")
    s.set_conserved(U)
    proj = M.ct_faces_from_cells(s)
    d0 = M.max_divb_face(s)
    el = timed_run(M, s, steps, "ref")
    out = {"workload": f"CT GLM-free step FP32 {n}^3 on the plain run's inputs, {steps} steps (fused :ref kernel)",
           "value": n ** 3 * steps / el, "projection": {k: float(v) for k, v in proj.items()},
           "max_divb_face_before": d0, "max_divb_face_after": M.max_divb_face(s),
           "finite": bool(all(np.isfinite(a).all() for a in s.fields_to_host()))}
    M.set_ct_kernel(s, True)
    k = max(steps // 10, 2)
    out["per_cell_kernel_value"] = n ** 3 * k / timed_run(M, s, k, "ref")
    return out


def coupled_leg(M, U, n, steps):
    """The drag and radiation-coupled fused steps (step_cube_drag!, step_cube_radiation!
    with density, field and energy tracking; integrator_cube.jl:478-540), which the
    production runner calls instead of step! (runner.jl:6660-6694), on the plain run's
    inputs with a fixed dt, drag impulse 0.1 and a synthetic velocity correction; psi is
    copied into the correction buffer before each radiation step, as the radiation driver
    does (radiation_coupling.jl:1681). kernel_ms is the fused step kernel alone."""
    import numpy as np
    s = M.allocate_state((n, n, n), 1.0 / n, dtype=np.float32, recon="plm_minmod")
    s.set_conserved(U)
    c = np.zeros((9, n ** 3), np.float32)
    c[1:4] = np.random.default_rng(1).normal(0, 0.01, (3, n ** 3)).astype(np.float32)
    buf = M.correction_buffer(s, c)
    dt, sm = M.compute_dt(s, 0.2)
    out = {"workload": f"coupled cube steps FP32 {n}^3 PLM minmod HLLD on the plain run's inputs, "
                       f"fixed CFL-0.2 dt, drag impulse 0.1, {steps} steps each"}

    def drag():
        M.step_drag(s, float(dt), float(sm), 0.99, 0.1, integrator="cube")

    def radiation():
        M.load_psi(s, buf)
        M.step_radiation(s, float(dt), float(sm), 0.99, 0.1, buf, integrator="cube")

    for name, fn in (("drag", drag), ("radiation", radiation)):
        fn()
        s.synchronize()
        s.timing(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        s.synchronize()
        el = time.perf_counter() - t0
        kms, nl = s.timing_read()
        s.timing(False)
        out[name] = {"value": n ** 3 * steps / el, "ms_per_step": el * 1e3 / steps,
                     "kernel_ms": kms / max(nl, 1)}
    out["finite"] = bool(all(np.isfinite(a).all() for a in s.fields_to_host()))
    return out


def refbench_leg(M, P, n, iters, solver):
    """The reference's own benchmark (scripts/benchmark.jl:41-72): float32, PPM, HLLD,
    the init_turb_field! state (problems.jl:9-29), one dt and smax from compute_dt at CFL
    0.2, then `iters` blocking steps with that dt (ch = smax; CT ignores it; each step waits
    for the GPU, as the reference's does) on the GPU
    default integrator (GLM: cube; CT: ref), N^3 iters / wall time. The recorded figures
    are the reference's only GPU throughput numbers (Apple M5 Max, Metal)."""
    import numpy as np
    s = M.allocate_state((n, n, n), 1.0 / n, dtype=np.float32, recon="ppm", riemann="hlld", mhd_solver=solver)
    U = P.turb_field((n, n, n), 1.0 / n)

    def init():
        s.set_conserved(U)
        if solver == "ct":
            M.ct_faces_from_cells(s)

    integ = "ref" if solver == "ct" else "cube"
    init()
    dt, sm = M.compute_dt(s, 0.2)

    def advance():
        # blocking like the reference's step! (its launchers end with KA.synchronize)
        if solver == "ct":
            M.step(s, float(dt), integrator=integ)
        else:
            M.step(s, float(dt), ch=float(sm), integrator=integ)
        s.synchronize()

    advance()
    init()  # the warm-up must not advance the measured state (benchmark.jl:50)
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        advance()
    s.synchronize()
    el = time.perf_counter() - t0
    h = s.fields_to_host()
    ref = {"glm": (111.05188196743987, "docs/data/benchmark-metal.toml:10 (N=128, 100 iterations)"),
           "ct": (133.1, "docs/performance.qmd:52-57 (CT ref, N=96, 60 iterations, mean of 131.9 and 134.3)")}
    mc = n ** 3 * iters / el / 1e6
    return {"workload": f"scripts/benchmark.jl --mhd-solver={solver}: float32 PPM HLLD {n}^3, init_turb_field!, "
                        f"fixed CFL-0.2 dt, {iters} steps, integrator {integ}",
            "mcells_per_second": mc, "reference_mcells_per_second": ref[solver][0],
            "reference_source": ref[solver][1], "reference_hardware": "Apple M5 Max GPU (Metal)",
            "ratio_to_reference": mc / ref[solver][0],
            "finite": bool(all(np.isfinite(a).all() for a in h)), "positive_density": bool(h[0].min() > 0)}


def valu_from_profile(n, integ, kernel_s):
    """VALU lane operations per cell-update from the committed SQ counter summary, and the
    issue rate they imply at the measured kernel time."""
    path = os.path.join(ROOT, "profiles", "k_step_fused_256cubed_sq_counters.md")
    if n != 256 or integ != "cube" or not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line and line.split(":")[0] in ("lane_ops_per_cell", "issue_slots_per_cell"):
            vals[line.split(":")[0]] = float(line.split(":")[1])
    if len(vals) < 2:
        return None
    ach = vals["issue_slots_per_cell"] * n ** 3 / kernel_s
    return {"bound": "valu-issue", "lane_ops_per_cell": vals["lane_ops_per_cell"],
            "issue_slots_per_cell": vals["issue_slots_per_cell"], "achieved": ach, "peak": VALU_PEAK_LANE_OPS,
            "unit": "lane-issue-slots/s", "frac": ach / VALU_PEAK_LANE_OPS,
            "note": "counted at 256^3: SQ_INSTS_VALU x 64 / cells, transcendentals at four slots; peak at the "
                    "2.4 GHz engine clock, which the chip does not hold under this load (2.13 GHz under the profiler)"}


def traffic_from_profile(n, integ):
    """HBM bytes per fused launch from the committed FETCH_SIZE / WRITE_SIZE summary."""
    path = os.path.join(ROOT, "profiles", "k_step_fused_256cubed_traffic.md")
    if n != 256 or integ != "cube" or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--integrator", default="auto", choices=["auto", "cube", "ref"])
    ap.add_argument("--cpu-steps", type=int, default=5,
                    help="steps in the CPU-baseline sample (0 = skip; 5 steps at 256^3 are about 12 s of CPU work)")
    ap.add_argument("--error-steps", type=int, default=200,
                    help="steps of the float32-vs-float64 error report (0 = skip)")
    ap.add_argument("--surrogate-steps", type=int, default=200,
                    help="steps of the surrogate-vs-plain leg (0 = skip)")
    ap.add_argument("--surrogate-weights", default=os.path.join(ROOT, "pmf-glm_amd", "pmfmhd", "data",
                                                                  "surrogate_weights.npy"))
    ap.add_argument("--no-ensemble", action="store_true", help="skip the ensemble leg")
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in z-chunk depth")
    ap.add_argument("--mach15-steps", type=int, default=100, help="steps of the 512^3 leg (0 = skip)")
    ap.add_argument("--fp64-oracle-steps", type=int, default=5,
                    help="steps of the float64 GPU-vs-CPU-oracle check at full size (0 = skip; ~3 s per step)")
    ap.add_argument("--oracle-error-n", type=int, default=128,
                    help="grid of the float32-GPU-vs-float64-CPU-oracle error report (0 = skip)")
    ap.add_argument("--oracle-error-steps", type=int, default=200,
                    help="steps of that report (about 75 s of oracle work at 128^3 on 16 cores)")
    ap.add_argument("--ct-steps", type=int, default=200, help="steps of the CT leg (0 = skip)")
    ap.add_argument("--coupled-steps", type=int, default=20,
                    help="steps of each coupled (drag, radiation) leg (0 = skip)")
    ap.add_argument("--refbench", action=argparse.BooleanOptionalAction, default=True,
                    help="the reference's own benchmark workload (scripts/benchmark.jl), GLM at N=128 and CT at N=96")
    ap.add_argument("--surrogate-profile", type=float, default=None, metavar="THR",
                    help="only run --steps surrogate-assisted steps at gate THR on the 256^3 surrogate inputs "
                         "(after --warmup), for a profiler; prints one JSON line of its own")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after timing, copy the 9 state fields device-to-device once (a plain copy of known "
                         "size that calibrates the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the GLM step is "
              "single-device; multi-GPU halo exchange is a design note only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import pmfmhd as M
    from pmfmhd import problems as P

    n = args.n
    if args.surrogate_profile is not None:
        print(json.dumps(surrogate_profile(M, P, n, args.steps, args.warmup, args.surrogate_weights,
                                           args.surrogate_profile)), flush=True)
        return
    log(f"bench: building the {n}^3 initial state")
    t0 = time.perf_counter()
    U = P.config_state("turb_256", seed=1, n=n)
    log(f"bench: initial state ready ({time.perf_counter() - t0:.1f} s)")
    s = M.allocate_state((n, n, n), 1.0 / n, dtype=np.float32, recon="plm_minmod", riemann="hlld")
    s.set_conserved(U)
    integ = args.integrator
    if integ == "auto":
        try:
            dt, sm = M.compute_dt(s, 0.2)
            M.step(s, float(dt), ch=float(sm), integrator="cube")
            integ = "cube"
        except M.MHDError as e:
            log(f"bench: fused kernel unavailable ({e}); using the per-cell kernel")
            integ = "ref"
        s.set_conserved(U)

    tuned = None
    if integ == "cube" and not args.no_autotune:
        # the tuner writes only the scratch fields: the state is still U, and the warm-up
        # follows on a GPU that has just been busy (no idle upload in between)
        zc, ms = M.autotune(s, reps=5)
        tuned = {"zc": zc, "ms": ms}
        log(f"bench: autotuned z chunk {zc} planes ({ms:.3f} ms/step)")

    def one_step():
        dt, sm = M.compute_dt(s, 0.2)
        M.step(s, float(dt), ch=float(sm), glm_cr=M.GLM_CR, integrator=integ)

    # The fused path steps with pmf_advance: every step's dt, ch and psi decay are computed
    # on the device from the speed the previous step folded (evolve!'s rule with
    # glm_ch_fac 1), bit for bit the compute_dt + step pair, without a host round trip.
    device_dt = integ == "cube"
    log(f"bench: warmup {args.warmup} steps ({integ})")
    if device_dt:
        M.advance(s, args.warmup, cfl=0.2)
    else:
        for _ in range(args.warmup):
            one_step()
    s.synchronize()
    s.timing(True)
    log(f"bench: timing {args.steps} steps")
    t0 = time.perf_counter()
    if device_dt:
        M.advance(s, args.steps, cfl=0.2)
    else:
        for _ in range(args.steps):
            one_step()
    s.synchronize()
    el = time.perf_counter() - t0
    kms, nl = s.timing_read()
    s.timing(False)
    host_loop = None
    if device_dt:
        # the same steps driven from the host (one compute_dt round trip per step)
        hn = min(args.steps, 20)
        s.synchronize()
        t1 = time.perf_counter()
        for _ in range(hn):
            one_step()
        s.synchronize()
        el1 = time.perf_counter() - t1
        host_loop = {"steps": hn, "ms_per_step": el1 / hn * 1e3, "value": n ** 3 * hn / el1}
    cells = n ** 3
    value = cells * args.steps / el
    avg_kernel_s = kms / 1e3 / max(nl, 1)
    achieved = BYTES_PER_CELL * cells / avg_kernel_s / 1e9
    h = s.fields_to_host()
    finite = all(np.isfinite(a).all() for a in h)
    log(f"bench: {value / 1e9:.3f} Gcell/s, step kernel {avg_kernel_s * 1e3:.3f} ms, finite={finite}")

    if args.calib_copy:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        src = s.device_ptr(0)
        dst = ctypes.c_void_p()
        M.lib().pmf_scratch_ptr(s.handle, 0, ctypes.byref(dst))
        nbytes = 9 * cells * 4
        rc = hip.hipMemcpy(dst, ctypes.c_void_p(src), nbytes, 3)  # hipMemcpyDeviceToDevice
        log(f"bench: calibration copy of {nbytes} bytes rc={rc}")

    cpu = None
    if args.cpu_steps > 0:
        log(f"bench: CPU baseline ({args.cpu_steps} step(s))")
        cpu = cpu_baseline(U, n, args.cpu_steps)

    del s
    errors = sur = ens = None
    if args.error_steps > 0 and integ == "cube":
        log(f"bench: float32 vs float64 error report ({args.error_steps} steps)")
        errors = error_report(M, P, n, args.error_steps)
    if args.surrogate_steps > 0 and integ == "cube" and os.path.exists(args.surrogate_weights):
        from pmfmhd.surrogate import THR
        log(f"bench: surrogate leg ({args.surrogate_steps} steps)")
        sur = surrogate_leg(M, P, n, args.surrogate_steps, np.load(args.surrogate_weights), THR)
    if not args.no_ensemble:
        log("bench: ensemble leg")
        ens = ensemble_leg(P)
    m15 = ct = orc = oerr = None
    if args.oracle_error_n > 0 and args.oracle_error_steps > 0 and integ == "cube":
        log(f"bench: float32 GPU vs float64 CPU oracle error report ({args.oracle_error_n}^3, "
            f"{args.oracle_error_steps} steps)")
        oerr = oracle_error_report(M, P, args.oracle_error_n, args.oracle_error_steps)
    if args.fp64_oracle_steps > 0 and integ == "cube":
        log(f"bench: float64 GPU vs CPU oracle ({args.fp64_oracle_steps} steps)")
        orc = fp64_oracle_check(M, P, n, args.fp64_oracle_steps)
    if args.mach15_steps > 0:
        log(f"bench: 512^3 leg ({args.mach15_steps} steps)")
        m15 = mach15_leg(M, P, args.mach15_steps)
    if args.ct_steps > 0:
        log(f"bench: CT leg ({args.ct_steps} steps)")
        ct = ct_leg(M, P, n, args.ct_steps)
    coupled = None
    if args.coupled_steps > 0 and integ == "cube":
        log(f"bench: coupled-step leg ({args.coupled_steps} steps each)")
        coupled = coupled_leg(M, U, n, args.coupled_steps)
        log("bench: coupled steps " + ", ".join(f"{k} {coupled[k]['ms_per_step']:.3f} ms/step "
                                                f"(kernel {coupled[k]['kernel_ms']:.3f} ms)"
                                                for k in ("drag", "radiation")))
    refb = None
    if args.refbench:
        log("bench: the reference's benchmark workload")
        refb = {"glm_n128": refbench_leg(M, P, 128, 100, "glm"), "ct_n96": refbench_leg(M, P, 96, 60, "ct")}

    kernel = "k_step_fused" if integ == "cube" else "k_step_ref"
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded solenoidal Gaussian B (P~k^2, |m|<=N/8, rms 1) and v (P~k^-2, Mach 0.5), "
                "rho = 1 + 0.01 N(0,1), p = 1",
        "config": {"workload": f"plain GLM-HLLD FP32 {n}^3, PLM minmod, CFL 0.2, gamma 5/3, periodic",
                   "n": n, "recon": "plm_minmod", "riemann": "hlld", "cfl": 0.2, "integrator": integ,
                   "parallelism": "single GPU", "finite": bool(finite), "autotuned_z_chunk": tuned,
                   "dt": "on the device (pmf_advance)" if device_dt else "on the host (compute_dt per step)"},
        "host_driven_steps": host_loop,
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profile(n, integ),
                     "kernel_ms": avg_kernel_s * 1e3, "bytes_per_launch": BYTES_PER_CELL * cells},
        "valu_roofline": valu_from_profile(n, integ, avg_kernel_s),
        "cpu_baseline": cpu,
        "error_report": dict(oerr or {}, float64_gpu_reference_at_benchmark_size=errors),
        "surrogate": sur,
        "ensemble": ens,
        "mach15_512": m15,
        "ct": ct,
        "fp64_oracle_check": orc,
        "coupled_steps": coupled,
        "reference_benchmark": refb,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
