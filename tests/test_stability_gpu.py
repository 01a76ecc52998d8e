"""Two reference checks run by the HIP library: the CFL ceilings of the unsplit scheme
(test/test_cfl_stability.jl with tools/cfl_stability/cfl_stability.jl: stable for 600
steps at the accepted ceiling, blown up before that at twice it) and planar invariance
after a step (test/streaming_checks.jl:50-68: a state invariant along one axis stays
invariant). Float64 for the stability scan, as the reference's harness."""
import numpy as np
import pytest

import pmfmhd as M
from pmfmhd import problems as P

pytestmark = pytest.mark.gpu
G = 5.0 / 3.0
CEILING = {"glm": 0.25, "ct": 0.30}  # MAX_STABLE_CFL (src/configuration.jl)


class _Stop(Exception):
    pass


def _seeded_state(N, eps0=1e-6, seed=11, B0=(0.8, 0.5, 0.4)):
    """cfl_stability.jl:14-40: rho = p = 1, B = B0 + centred curl of a 1e-6 noise
    potential, an oblique k = (1,2,3) wave plus 1e-6 white noise (numpy's generator, so
    the noise realisation differs from the reference's)."""
    rng = np.random.default_rng(seed)
    k, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    ph = 2 * np.pi * ((i + 1) + 2 * (j + 1) + 3 * (k + 1)) / N
    A = [eps0 * rng.standard_normal((N, N, N)) for _ in range(3)]

    def d(a, axis):  # centred difference along axis (0 = z, 1 = y, 2 = x) times N/2
        return (np.roll(a, -1, axis) - np.roll(a, 1, axis)) * N / 2

    bx = B0[0] + d(A[2], 1) - d(A[1], 0)
    by = B0[1] + d(A[0], 0) - d(A[2], 2)
    bz = B0[2] + d(A[1], 2) - d(A[0], 1)
    rho = 1 + eps0 * (np.sin(ph) + rng.standard_normal((N, N, N)))
    vx = eps0 * (0.5 * np.cos(ph) + rng.standard_normal((N, N, N)))
    vy = eps0 * (0.3 * np.cos(ph) + rng.standard_normal((N, N, N)))
    vz = eps0 * (0.2 * np.cos(ph) + rng.standard_normal((N, N, N)))
    p = 1 + eps0 * np.sin(ph)
    U = np.zeros((9, N, N, N))
    U[0], U[1], U[2], U[3] = rho, rho * vx, rho * vy, rho * vz
    U[5], U[6], U[7] = bx, by, bz
    U[4] = p / (G - 1) + 0.5 * rho * (vx ** 2 + vy ** 2 + vz ** 2) + 0.5 * (bx ** 2 + by ** 2 + bz ** 2)
    return U.reshape(9, -1)


def _stability_run(solver, cfl, integrator, nsteps=600, N=16):
    s = M.allocate_state((N, N, N), 1.0 / N, dtype=np.float64, recon="plm", mhd_solver=solver)
    s.set_conserved(_seeded_state(N))
    if solver == "ct":
        M.ct_faces_from_cells(s)
    amp = lambda: float(np.abs(s.download(0) - 1.0).max())
    a0 = amp()
    rec = {"steps": 0, "amax": a0, "blown": False}

    def cb(st, t, n):
        a = amp()
        rec["steps"] = n
        rec["amax"] = max(rec["amax"], a)
        if not np.isfinite(a) or a > 0.1:
            rec["blown"] = True
            raise _Stop()
        if n >= nsteps:
            raise _Stop()

    try:
        M.evolve(s, 1e9, cfl=cfl, glm_ch_fac=2.0, integrator=integrator, callback=cb)
    except _Stop:
        pass
    rec["growth"] = rec["amax"] / a0
    return rec


@pytest.mark.parametrize("solver,integrator", [("glm", "ref"), ("glm", "cube"), ("ct", "ref")])
def test_cfl_ceiling_is_stable_and_twice_it_is_not(solver, integrator):
    ok = _stability_run(solver, CEILING[solver], integrator)
    assert not ok["blown"] and ok["steps"] == 600
    assert ok["growth"] < 100  # physical wave plus noise (the reference measured ~4 GLM, ~10 CT)
    bad = _stability_run(solver, 2 * CEILING[solver], integrator)
    assert bad["blown"] and bad["steps"] < 600


@pytest.mark.parametrize("integrator", ["ref", "cube"])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_planar_invariance_after_a_step(axis, integrator):
    """A turbulent state averaged along one axis (so invariant along it, B still
    solenoidal), boosted by (0.1, 0.2, 0.3), stays invariant along that axis after a step:
    every cell of a line sees the same arithmetic, so exactly."""
    N = 32
    U = P.config_state("turb_256", seed=5, n=N).reshape(9, N, N, N)
    ax = 3 - axis  # array axes are (var, z, y, x); axis 0 = x
    U = np.broadcast_to(U.mean(axis=ax, keepdims=True), U.shape).copy()
    rho = U[0]
    ek0 = 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / rho
    for c, v in enumerate((0.1, 0.2, 0.3)):
        U[1 + c] += rho * v
    U[4] += 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / rho - ek0
    U = U.reshape(9, -1).astype(np.float32)
    s = M.allocate_state((N, N, N), 1.0 / N, dtype=np.float32, recon="plm")
    s.set_conserved(U)
    M.step(s, 1e-4, ch=1.0, integrator=integrator)
    out = s.conserved().reshape(9, N, N, N)
    shift = np.roll(out, 1, axis=ax)
    assert np.array_equal(out, shift)
    assert np.abs(out[5:8]).max() > 1e-3
