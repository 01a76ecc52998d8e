"""The 512^3 benchmark configuration's recipe (Mach-1.5 solenoidal velocity with
P(k)~k^-2, lognormal rho with sigma 0.3, solenoidal B with P(k)~k^2 and rms 1, |m| <= N/8)
at 32^3: the workload where shocks form, so HLLD's four flux branches, its degenerate
single-star guard and the LLF switch are all reached. The census of which faces take
which branch is taken on the CPU from the numpy restatement's face states; the GPU kernels
are then checked against the oracle on the same states."""
import numpy as np
import pytest

import face_np as F
import oracle as O
from pmfmhd import problems as P

G = 5.0 / 3.0
N = 32


def mach15_state(seed=2, n=N):
    return P.config_state("mach15_512", seed=seed, n=n)


def branch_census(U, n, dt, ch, llf_dmin=1e-4, recon="plm_minmod"):
    """Counts of faces per HLLD dispatch outcome (-1 = LLF switch, 0..3 = flux branch)
    over the three face sets of one step, and the number of degenerate single stars."""
    counts = np.zeros(5, dtype=np.int64)
    degenerate = [0]

    def hook(L, R, d):
        br, dg = F.hlld_branches(L, R, d, ch, G, llf_dmin=llf_dmin)
        counts[:] += np.bincount(br.ravel() + 1, minlength=5)
        degenerate[0] += int(dg.sum())
        return F.riemann(L, R, d, ch, G, llf_dmin=llf_dmin)

    dx = 1.0 / n
    F.step_periodic(U, (n, n, n), dt / dx, ch, 1.0, llf_dmin=llf_dmin, recon=recon, flux=hook)
    return dict(llf=int(counts[0]), branches=[int(c) for c in counts[1:]], degenerate=degenerate[0])


def _dt(U, n, llf_dmin):
    o = O.OracleState((n, n, n), 1.0 / n, recon="plm_minmod", llf_dmin=llf_dmin)
    o.set_conserved(U)
    return o.compute_dt(0.2)


def test_mach15_faces_reach_every_hlld_branch_and_the_llf_switch():
    U = mach15_state()
    rho = U[0]
    assert rho.min() > 0.2 and abs(rho.mean() - 1) < 0.05  # lognormal, unit mean
    dt, smax = _dt(U, N, 1e-4)
    c = branch_census(U, N, dt, smax)
    assert c["llf"] == 0  # no density below the default llf_dmin = 1e-4
    assert min(c["branches"]) > 100, c
    # a raised llf_dmin hands the low-density faces to LLF and leaves HLLD the others
    c = branch_census(U, N, dt, smax, llf_dmin=0.8)
    assert c["llf"] > 1000 and min(c["branches"]) > 100, c


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", ["cube", "ref"])
@pytest.mark.parametrize("llf_dmin", [1e-4, 0.8])
def test_mach15_float64_ten_steps_match_oracle(integrator, llf_dmin):
    """float64 HIP against the float64 oracle to round-off over 10 CFL steps on the
    shock-forming workload, with the default LLF switch and with one that takes every face
    next to a density below 0.8."""
    import pmfmhd as M
    U = mach15_state()
    kw = dict(recon="plm_minmod", llf_dmin=llf_dmin)
    g = M.allocate_state((N,) * 3, 1.0 / N, dtype=np.float64, **kw)
    o = O.OracleState((N,) * 3, 1.0 / N, **kw)
    g.set_conserved(U)
    o.set_conserved(U)
    for _ in range(10):
        dt, smax = o.compute_dt(0.2)
        _, gsmax = M.compute_dt(g, 0.2)
        assert abs(gsmax - smax) <= 1e-13 * smax
        o.step(float(dt), ch=float(smax))
        M.step(g, float(dt), ch=float(smax), integrator=integrator)
    Ug = g.conserved()
    scale = np.abs(o.U).max(axis=1, keepdims=True)
    assert np.max(np.abs(Ug - o.U) / scale) < 1e-12
    assert np.all(Ug[0] > 0)


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", ["cube", "ref"])
def test_mach15_float32_one_step_matches_float32_oracle(integrator):
    import pmfmhd as M
    U = mach15_state().astype(np.float32)
    g = M.allocate_state((N,) * 3, 1.0 / N, dtype=np.float32, recon="plm_minmod")
    o = O.OracleState((N,) * 3, 1.0 / N, dtype=np.float32, recon="plm_minmod")
    g.set_conserved(U)
    o.set_conserved(U)
    dt, smax = o.compute_dt(0.2)
    o.step(float(dt), ch=float(smax))
    M.step(g, float(dt), ch=float(smax), integrator=integrator)
    Ug = g.conserved().astype(np.float64)
    for v in range(9):
        ref = o.U[v].astype(np.float64)
        if v == 8:  # psi starts at zero: absolute, against the field scale
            assert np.linalg.norm(Ug[v] - ref) / np.linalg.norm(U[5:8]) < 1e-5
        else:
            assert np.linalg.norm(Ug[v] - ref) / np.linalg.norm(ref) < 1e-5, v
