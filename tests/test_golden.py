"""The committed float64 fixtures (tests/golden, written by make_golden.py from the CPU
oracle): the oracle must still reproduce them, and the HIP library must match them."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
FACES = dict(np.load(os.path.join(HERE, "face_fluxes.npz")))
STEPS = dict(np.load(os.path.join(HERE, "steps.npz")))


def _rel(a, b):
    scale = np.maximum(np.abs(b).max(axis=-1, keepdims=True), 1e-30)
    return float((np.abs(np.asarray(a) - b) / scale).max())


def test_oracle_reproduces_the_fixtures():
    import make_golden
    faces, steps = make_golden.compute()
    assert set(faces) == set(FACES) and set(steps) == set(STEPS)
    for k in FACES:
        assert _rel(faces[k], FACES[k]) < 1e-14, k
    for k in STEPS:
        assert _rel(steps[k], STEPS[k]) < 1e-13, k


@pytest.mark.gpu
def test_library_face_fluxes_match_fixtures():
    import pmfmhd as M
    L, R = FACES["L"], FACES["R"]
    for rs in ("hlld", "hll", "llf"):
        s = M.allocate_state((4, 4, 4), 0.25, dtype=np.float64, riemann=rs)
        for d in range(3):
            assert _rel(M.riemann_flux(s, L, R, d, ch=1.3), FACES[f"F_{rs}_{d}"]) < 1e-13, (rs, d)
    c = M.allocate_state((4, 4, 4), 0.25, dtype=np.float64, riemann="hll", mhd_solver="ct")
    for d in range(3):
        assert _rel(M.riemann_flux(c, L, R, d), FACES[f"F_ct_{d}"]) < 1e-13, d


@pytest.mark.gpu
@pytest.mark.parametrize("name,dims,recon,rs,bcs,integ", [
    ("plm_hlld", (6, 6, 6), "plm", "hlld", ("periodic",) * 3, "ref"),
    ("ppm_hll", (6, 6, 6), "ppm", "hll", ("periodic",) * 3, "ref"),
    ("minmod_llf", (6, 6, 6), "plm_minmod", "llf", ("periodic",) * 3, "ref"),
    ("plm_hlld_bcs", (6, 6, 6), "plm", "hlld", ("outflow", "reflecting", "periodic"), "ref"),
    ("fused_minmod_hlld", (16, 12, 2), "plm_minmod", "hlld", ("periodic",) * 3, "cube"),
    ("fused_minmod_hlld", (16, 12, 2), "plm_minmod", "hlld", ("periodic",) * 3, "ref"),
])
def test_library_steps_match_fixtures(name, dims, recon, rs, bcs, integ):
    import pmfmhd as M
    s = M.allocate_state(dims, 1.0 / max(dims), dtype=np.float64, recon=recon, riemann=rs, bcs=bcs)
    s.set_conserved(STEPS[f"{name}/U0"])
    dt, ch = STEPS[f"{name}/dt_ch"]
    M.step(s, dt, ch=ch, integrator=integ)
    assert _rel(s.conserved(), STEPS[f"{name}/U1"]) < 1e-12


@pytest.mark.gpu
def test_library_ct_and_coupled_steps_match_fixtures():
    import pmfmhd as M
    c = M.allocate_state((6, 6, 6), 1 / 6, dtype=np.float64, recon="plm", riemann="hll", mhd_solver="ct")
    c.set_conserved(STEPS["ct/U0"])
    M.ct_install_faces(c, *STEPS["ct/faces0"])
    M.step(c, float(STEPS["ct/dt"][0]))
    assert _rel(c.conserved(), STEPS["ct/U1"]) < 1e-12
    for d in range(3):
        assert _rel(c.ct_slot(d), STEPS["ct/faces1"][d]) < 1e-12
    s = M.allocate_state((6, 6, 6), 1 / 6, dtype=np.float64, recon="plm", riemann="hlld")
    s.set_conserved(STEPS["drag/U0"])
    dt, ch = STEPS["drag/dt_ch"]
    M.step_drag(s, dt, ch, 0.93, 0.4, target_velocity=(0.1, -0.2, 0.05))
    assert _rel(s.conserved(), STEPS["drag/U1"]) < 1e-12
    s.set_conserved(STEPS["radiation/U0"])
    buf = M.correction_buffer(s, STEPS["radiation/corr0"])
    dt, ch = STEPS["radiation/dt_ch"]
    M.step_radiation(s, dt, ch, 0.93, 0.3, buf)
    assert _rel(s.conserved(), STEPS["radiation/U1"]) < 1e-12
    got = buf.download(np.float64, (9, 216))
    assert np.abs(got - STEPS["radiation/corr1"]).max() < 1e-12
