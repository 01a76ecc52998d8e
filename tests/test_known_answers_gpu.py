"""The reference's recorded known answers (validation-cpu.log, transcribed in
tests/golden/known_answers.json) reproduced by the HIP library: Alfvén-wave convergence
(PLM, PPM), low-beta Alfvén, Brio-Wu, Orszag-Tang and field-loop advection
(test/inherited/runtests.jl:6-122,486-509), on the per-cell and on the fused integrator.
Float32, the reference's own precision for these records."""
import json
import os

import numpy as np
import pytest

import pmfmhd as M
from pmfmhd import problems as P

pytestmark = pytest.mark.gpu
KNOWN = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "known_answers.json")))


def _alfven(N, recon, integrator, amp=0.1, p0=0.1, glm_cr=M.GLM_CR):
    s = M.allocate_state((N, 4, 4), 1.0 / N, dtype=np.float32, recon=recon)
    s.set_conserved(P.alfven_wave((N, 4, 4), 1.0 / N, amp=amp, B0=1.0, p0=p0))
    t0 = M.conserved_totals(s)
    t, n = M.evolve(s, 0.5, cfl=0.4, glm_cr=glm_cr, integrator=integrator)
    t1 = M.conserved_totals(s)
    x = (np.arange(N) + 0.5) / N
    L1 = np.mean(np.abs(s.download(6)[:N].astype(np.float64) - P.alfven_By_exact(x, t, 1.0, amp=amp)))
    return L1, n, abs(t1["mass"] - t0["mass"]) / t0["mass"], abs(t1["energy"] - t0["energy"]) / t0["energy"]


@pytest.mark.parametrize("integrator", ["ref", "cube"])
@pytest.mark.parametrize("recon", ["plm", "ppm"])
def test_alfven_convergence(recon, integrator):
    """runtests.jl:6-62; validation-cpu.log:8-19."""
    L1s = []
    for rec, N in zip(KNOWN["alfven_circular"][recon], (32, 64, 128)):
        L1, n, dm, de = _alfven(N, recon, integrator)
        assert n == rec["steps"], (N, n)
        # float32 round-off (IEEE here, @fastmath in the reference): ~0.1% at N <= 64, ~2%
        # at N = 128 where the discretisation error is small
        assert abs(L1 - rec["L1"]) / rec["L1"] < (5e-3 if N < 128 else 2.5e-2), (N, L1, rec["L1"])
        assert dm < 1e-4 and de < 1e-4
        L1s.append(L1)
    assert min(np.log2(L1s[0] / L1s[1]), np.log2(L1s[1] / L1s[2])) > 1.7


@pytest.mark.parametrize("integrator", ["ref", "cube"])
def test_low_beta_alfven(integrator):
    """runtests.jl:27-35; validation-cpu.log:5."""
    K = KNOWN["alfven_low_beta"]
    L1, n, _, de = _alfven(K["N"], "plm", integrator, amp=0.1, p0=1e-3, glm_cr=1.0)
    assert n == K["steps"]
    assert abs(L1 - K["L1"]) / K["L1"] < 1e-2
    assert de < 2e-4


@pytest.mark.parametrize("integrator", ["ref", "cube"])
def test_brio_wu(integrator):
    """runtests.jl:65-91; validation-cpu.log:22 (outflow x boundaries)."""
    def density(N):
        s = M.allocate_state((N, 4, 4), 1.0 / N, dtype=np.float32, gamma=2.0,
                             bcs=("outflow", "periodic", "periodic"))
        s.set_conserved(P.brio_wu((N, 4, 4), gamma=2.0))
        t0 = M.conserved_totals(s)
        M.evolve(s, 0.1, cfl=0.4, integrator=integrator)
        t1 = M.conserved_totals(s)
        return s.download(0)[:N].astype(np.float64), abs(t1["mass"] - t0["mass"]) / t0["mass"]
    r2, _ = density(256)
    r4, dm4 = density(512)
    L1 = np.mean(np.abs(r2 - 0.5 * (r4[0::2] + r4[1::2])))
    K = KNOWN["brio_wu"]
    assert abs(L1 - K["L1_256_vs_512"]) / K["L1_256_vs_512"] < 2e-2
    assert abs(r2.min() - K["rho_min"]) < 1.5e-3
    assert r2[0] > 0.9 and r2[-1] < 0.2 and r2.max() < 1.02
    assert dm4 < 1e-4


@pytest.mark.parametrize("integrator", ["ref", "cube"])
def test_orszag_tang(integrator):
    """runtests.jl:94-122; validation-cpu.log:25 (128^2 x 1)."""
    N = 128
    s = M.allocate_state((N, N, 1), 1.0 / N, dtype=np.float32)
    s.set_conserved(P.orszag_tang((N, N, 1), 1.0 / N))
    t0 = M.conserved_totals(s)
    t, n = M.evolve(s, 0.5, cfl=0.4, integrator=integrator)
    t1 = M.conserved_totals(s)
    h = np.stack(s.fields_to_host()).astype(np.float64)
    brms = np.sqrt(np.mean(h[5] ** 2 + h[6] ** 2 + h[7] ** 2))
    norm = M.max_divb(s) * (1.0 / N) / brms
    assert n == KNOWN["orszag_tang"]["steps"]
    # the reference's check is max|div B| dx / Brms < 5; the recorded 2.79 is a pointwise
    # maximum in a turbulent field, which round-off moves by a few percent (2.74-2.84 seen)
    assert norm < 5
    assert abs(norm - KNOWN["orszag_tang"]["max_divb_dx_over_brms"]) < 0.1
    assert abs(t1["mass"] - t0["mass"]) / t0["mass"] < 1e-4
    assert h[0].min() > 0


@pytest.mark.parametrize("integrator", ["ref", "cube"])
def test_field_loop(integrator):
    """runtests.jl:486-509; validation-cpu.log:86."""
    N = 64
    s = M.allocate_state((N, N, 4), 1.0 / N, dtype=np.float32)
    s.set_conserved(P.field_loop((N, N, 4), 1.0 / N))
    h0 = np.stack(s.fields_to_host()).astype(np.float64)
    bnorm = np.abs(h0[5]).sum() + np.abs(h0[6]).sum()
    emag0 = (h0[5] ** 2).sum() + (h0[6] ** 2).sum()
    assert M.max_divb(s) < 2e-6
    t, n = M.evolve(s, 1.0, cfl=0.4, glm_ch_fac=2.0, glm_cr=1.0, integrator=integrator)
    h1 = np.stack(s.fields_to_host()).astype(np.float64)
    l1 = (np.abs(h1[5] - h0[5]).sum() + np.abs(h1[6] - h0[6]).sum()) / bnorm
    emag1 = (h1[5] ** 2).sum() + (h1[6] ** 2).sum()
    divn = M.max_divb(s) * (1.0 / N) / np.sqrt(emag1 / h1[5].size)
    K = KNOWN["field_loop"]
    assert n == K["steps"]
    assert abs(l1 - K["L1"]) < 2e-3
    assert abs(emag1 / emag0 - K["emag_retention"]) < 2e-3
    assert divn < 0.5
