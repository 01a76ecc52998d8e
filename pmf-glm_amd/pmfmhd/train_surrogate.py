"""Train the flux surrogate on face-state / flux pairs dumped by the plain solver.

Data: plain fused-kernel runs of the 32^3 benchmark inputs and of the 256^3 turbulence
inputs over seeds 1-16 (each face that passes the jump gate is a candidate; a seeded
keep probability spreads the 2^24 pairs over the runs' steps). Seed 17 of both is held
out. Training: Adam on the GPU with the project's own forward/backward kernels
(bfloat16 MFMA, float32 accumulation and master weights), stepped learning rate.
Writes the float32 weights and a JSON report (dataset composition, loss history,
held-out per-face errors against HLLD).

    python -m pmfmhd.train_surrogate --out DIR [--pairs 16777216]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import allocate_state, compute_dt, step
from . import problems as P
from .surrogate import THR, Surrogate

LEARNED = [0, 1, 2, 3, 4, 6, 7]


def _log(msg):
    print(msg, flush=True)


def dump_runs(net, name, seeds, target, steps, cfl=0.2, thr=THR, log=_log):
    """Evolve each seed of configuration `name` with the plain fused step for `steps`
    steps, dumping gated faces before every step with a keep probability adapted so
    that about `target` pairs come from this configuration."""
    n = P.CONFIGS[name]["n"]
    s = allocate_state((n, n, n), 1 / n, dtype=np.float32, recon="plm_minmod")
    start = net.count
    per_dump = target / (len(seeds) * steps)
    # first dump: a keep probability that cannot overflow the target even if every face
    # passes the gate; later dumps use the gated count it reveals
    keep0 = float(min(1.0, per_dump / (3 * n ** 3)))
    keep = None
    for seed in seeds:
        s.set_conserved(P.config_state(name, seed=seed))
        for it in range(steps):
            dt, smax = compute_dt(s, cfl)
            before = net.count
            k = keep0 if keep is None else keep
            net.dump(s, float(dt), float(smax), thr=thr, keep=k, seed=1000003 * seed + it)
            if keep is None:
                gated = max((net.count - before) / keep0, 1.0)
                keep = float(min(1.0, per_dump / gated))
            step(s, float(dt), ch=float(smax), integrator="cube")
        log(f"  {name} seed {seed}: {net.count - start} pairs so far (keep {keep:.4g})")
    return net.count - start


def face_errors(net, first, count, thr=THR):
    """Held-out per-face errors on the learned slots: relative to |F_HLLD| and to the
    central flux's error (the method without the network)."""
    Fh = net.face_flux("hlld", first, count, thr)[:, LEARNED].astype(np.float64)
    Fs = net.face_flux("surrogate", first, count, thr)[:, LEARNED].astype(np.float64)
    Fc = net.face_flux("central", first, count, thr)[:, LEARNED].astype(np.float64)
    es = np.linalg.norm(Fs - Fh, axis=1)
    ec = np.linalg.norm(Fc - Fh, axis=1)
    fh = np.linalg.norm(Fh, axis=1)
    rel = es / np.maximum(fh, 1e-30)
    return {
        "faces": int(count),
        "rel_flux_error_median": float(np.median(rel)),
        "rel_flux_error_p99": float(np.quantile(rel, 0.99)),
        "rel_flux_error_max": float(rel.max()),
        "surrogate_over_central_error_l2": float(np.linalg.norm(es) / max(np.linalg.norm(ec), 1e-30)),
        "central_rel_flux_error_median": float(np.median(ec / np.maximum(fh, 1e-30))),
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--pairs", type=int, default=1 << 24)
    ap.add_argument("--seeds", type=int, default=16)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--schedule", default="6000:2e-3,6000:1e-3,6000:3e-4,6000:1e-4")
    ap.add_argument("--net-seed", type=int, default=2024)
    args = ap.parse_args(argv)
    os.makedirs(args.out, exist_ok=True)
    t0 = time.time()
    host = allocate_state((32, 32, 32), 1 / 32, dtype=np.float32)  # parameters and stream for the network
    net = Surrogate(host, seed=args.net_seed)
    net.reserve(args.pairs)
    seeds = list(range(1, args.seeds + 1))
    half = args.pairs // 2
    comp = {}
    _log("dumping training pairs")
    comp["bstatic_32"] = dump_runs(net, "bstatic_32", seeds, half, steps=64)
    comp["turb_256"] = dump_runs(net, "turb_256", seeds, args.pairs - net.count, steps=4)
    _log(f"dataset: {net.count} pairs {comp} ({time.time() - t0:.0f} s)")
    losses = []
    for phase in args.schedule.split(","):
        nsteps, lr = phase.split(":")
        t1 = time.time()
        l = net.train(int(nsteps), batch=args.batch, lr=float(lr), seed=len(losses) + 1)
        losses.extend(float(x) for x in l)
        _log(f"train {nsteps} steps at lr {lr}: loss {np.mean(l[:50]):.4g} -> {np.mean(l[-200:]):.4g} "
             f"({time.time() - t1:.1f} s)")
    w = net.weights()
    # held-out seed 17 of both configurations
    val = Surrogate(host, seed=1)
    val.set_weights(w)
    val.reserve(1 << 21)
    dump_runs(val, "bstatic_32", [17], 1 << 20, steps=16)
    n1 = val.count
    dump_runs(val, "turb_256", [17], 1 << 20, steps=2)
    report = {
        "pairs": int(net.count),
        "composition": comp,
        "threshold": THR,
        "batch": args.batch,
        "schedule": args.schedule,
        "final_loss": float(np.mean(losses[-500:])),
        "loss_every_100": [float(np.mean(losses[i:i + 100])) for i in range(0, len(losses), 100)],
        "heldout_bstatic_32": face_errors(val, 0, n1),
        "heldout_turb_256": face_errors(val, n1, val.count - n1),
        "seconds": time.time() - t0,
    }
    np.save(os.path.join(args.out, "surrogate_weights.npy"), w)
    with open(os.path.join(args.out, "surrogate_training.json"), "w") as f:
        json.dump(report, f, indent=1)
    _log(json.dumps({k: v for k, v in report.items() if k != "loss_every_100"}, indent=1))


if __name__ == "__main__":
    sys.exit(main())
