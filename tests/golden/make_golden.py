"""Writes the committed float64 fixtures of this directory from the CPU oracle.

    python tests/golden/make_golden.py

face_fluxes.npz   random normal-frame face pairs (floors and the LLF switch included) and
                  their HLLD / HLL / LLF / CT fluxes (glm_physics.jl:234-348)
steps.npz         one step of a smooth 6^3 state for several recon x riemann x boundary
                  cases, a CT step and the drag / radiation-coupled steps
                  (integrator_ref.jl:69-258, ct.jl:57-173), and one fused-kernel-shaped
                  16x12x2 periodic case
The reference ships no value fixtures for the step (it is Julia and checks itself
against its own CPU run); these pin the oracle against drift and give the GPU tests a
fixed target. Its recorded known answers are in known_answers.json.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "pmf-glm_amd"), os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

import oracle as O  # noqa: E402
from test_oracle import ct_test_hydro, ct_test_potential, smooth_state  # noqa: E402

G = 5.0 / 3.0

CASES = [  # name, dims, recon, riemann, bcs
    ("plm_hlld", (6, 6, 6), "plm", "hlld", ("periodic",) * 3),
    ("ppm_hll", (6, 6, 6), "ppm", "hll", ("periodic",) * 3),
    ("minmod_llf", (6, 6, 6), "plm_minmod", "llf", ("periodic",) * 3),
    ("plm_hlld_bcs", (6, 6, 6), "plm", "hlld", ("outflow", "reflecting", "periodic")),
    ("fused_minmod_hlld", (16, 12, 2), "plm_minmod", "hlld", ("periodic",) * 3),
]


def smooth(dims, seed):
    n = max(dims)
    U = smooth_state(n, seed=seed).reshape(9, n, n, n)
    nx, ny, nz = dims
    return np.ascontiguousarray(U[:, :nz, :ny, :nx].reshape(9, -1))


def faces(rng, n):
    q = np.empty((n, 9))
    q[:, 0] = rng.uniform(0.3, 2.0, n)
    q[:, 1:4] = rng.normal(0, 0.6, (n, 3))
    q[:, 4] = rng.uniform(0.1, 2.0, n)
    q[:, 5:8] = rng.normal(0, 0.8, (n, 3))
    q[:, 8] = rng.normal(0, 0.05, n)
    q[: n // 20, 0] = 1e-5  # below llf_dmin: the LLF switch
    q[n // 20: n // 10, 4] = 1e-9  # pressure floor
    return q


def compute():
    """(face fixtures, step fixtures) as dicts of arrays."""
    out = {}
    rng = np.random.default_rng(2024)
    L, R = faces(rng, 96), faces(rng, 96)
    out_f = {"L": L, "R": R}
    for d in range(3):
        for rs in ("hlld", "hll", "llf"):
            out_f[f"F_{rs}_{d}"] = O.riemann_batch(L, R, d, 1.3, riemann=rs)
        out_f[f"F_ct_{d}"] = O.riemann_batch(L, R, d, 0.0, ct=True, riemann="hll")

    for name, dims, recon, rs, bcs in CASES:
        U = smooth(dims, seed=len(name))
        s = O.OracleState(dims, 1.0 / max(dims), recon=recon, riemann=rs, bcs=bcs)
        s.set_conserved(U)
        dt, ch = s.compute_dt(0.3)
        s.step(dt, ch=ch)
        out[f"{name}/U0"] = U
        out[f"{name}/U1"] = s.U.copy()
        out[f"{name}/dt_ch"] = np.array([dt, ch])
    # CT (HLL, PLM MonCen, the fork-port potential field)
    N = 6
    s = O.OracleState((N, N, N), 1.0 / N, recon="plm", riemann="hll", mhd_solver="ct")
    s.set_conserved(ct_test_hydro(N, G))
    s.ct_faces_from_potential(*ct_test_potential(N))
    out["ct/U0"] = s.U.copy()
    out["ct/faces0"] = s.ct[:3].copy()
    dt, _ = s.compute_dt(0.3)
    s.step(dt)
    out["ct/U1"] = s.U.copy()
    out["ct/faces1"] = s.ct[:3].copy()
    out["ct/dt"] = np.array([dt])
    # coupled steps on the plm_hlld case: drag toward a moving target, radiation with tracking
    U = smooth((6, 6, 6), seed=3)
    for name, kw in (("drag", dict(drag_impulse=0.4, target_velocity=(0.1, -0.2, 0.05))),
                     ("radiation", dict(drag_impulse=0.3, track_density=True, track_magnetic=True,
                                        track_energy=True))):
        s = O.OracleState((6, 6, 6), 1.0 / 6, recon="plm", riemann="hlld")
        s.set_conserved(U)
        dt, ch = s.compute_dt(0.3)
        if name == "radiation":
            c = np.zeros((9, 216))
            c[0] = U[8]
            c[1:4] = np.random.default_rng(7).normal(0, 0.02, (3, 216))
            out["radiation/corr0"] = c.copy()
            kw["correction"] = c
        s.step_sources(dt, ch, 0.93, **kw)
        out[f"{name}/U0"] = U
        out[f"{name}/U1"] = s.U.copy()
        out[f"{name}/dt_ch"] = np.array([dt, ch])
        if name == "radiation":
            out["radiation/corr1"] = c.copy()
    return out_f, out


def main():
    out_f, out = compute()
    np.savez_compressed(os.path.join(HERE, "face_fluxes.npz"), **out_f)
    np.savez_compressed(os.path.join(HERE, "steps.npz"), **out)


if __name__ == "__main__":
    main()
