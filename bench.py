"""Throughput of the plain GLM-HLLD step on one MI355X.

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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--integrator", default="auto", choices=["auto", "cube", "ref"])
    ap.add_argument("--cpu-steps", type=int, default=1, help="steps in the CPU-baseline sample (0 = skip)")
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
    log(f"bench: building the {n}^3 initial state")
    t0 = time.perf_counter()
    U = P.config_state("c2_256", seed=1, n=n)
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

    def one_step():
        dt, sm = M.compute_dt(s, 0.2)
        M.step(s, float(dt), ch=float(sm), glm_cr=M.GLM_CR, integrator=integ)

    log(f"bench: warmup {args.warmup} steps ({integ})")
    for _ in range(args.warmup):
        one_step()
    s.synchronize()
    s.timing(True)
    log(f"bench: timing {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one_step()
    s.synchronize()
    el = time.perf_counter() - t0
    kms, nl = s.timing_read()
    s.timing(False)
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
                   "parallelism": "single GPU", "finite": bool(finite)},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel_ms": avg_kernel_s * 1e3, "bytes_per_launch": BYTES_PER_CELL * cells},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
