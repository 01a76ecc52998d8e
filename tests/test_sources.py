"""Drag / radiation-coupled steps (step_ref_drag!, step_ref_radiation!, step_cube_drag!,
step_cube_radiation!; integrator_ref.jl:28-48,150-160,222-258, integrator_cube.jl:68-116,
478-540, glm_physics.jl:63-170,350-390): pins of the oracle's source terms on the CPU,
then the GPU kernels against it."""
import numpy as np
import pytest

import oracle as O
from test_oracle import smooth_state

G = 5.0 / 3.0


def _uniform(n, rho=1.3, v=(0.4, -0.2, 0.3), p=0.8, B=(0.5, 0.2, -0.1), psi=0.0):
    q = np.zeros((9, n ** 3))
    q[0], q[4] = rho, p
    q[1:4] = np.array(v)[:, None]
    q[5:8] = np.array(B)[:, None]
    q[8] = psi
    from pmfmhd import problems as P
    return P.prim2cons(q, G)


def _eint(U):
    return U[4] - 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / U[0] - 0.5 * (U[5] ** 2 + U[6] ** 2 + U[7] ** 2)


def test_drag_coefficients_closed_forms_and_series_branches():
    for q in (0.3, 1.7, 0.06):
        e, phi1, A, C, D = O.drag_coefficients(q)
        em1, em2 = -np.expm1(-q), -np.expm1(-2 * q)
        assert e == pytest.approx(np.exp(-q), rel=1e-15)
        assert phi1 == pytest.approx(em1 / q, rel=1e-14)
        assert A == pytest.approx(0.5 * em2, rel=1e-14)
        assert C == pytest.approx(em1 * em1 / q, rel=1e-14)
        assert D == pytest.approx((q - 2 * em1 + 0.5 * em2) / q ** 2, rel=1e-9)
    # series branches (q < 1e-4 for phi1, q < 0.05 for A, C, D) against the exact forms
    # evaluated in long double-free ways: the truncation is O(q^3) / O(q^6)
    q = 9e-5
    assert O.drag_coefficients(q)[1] == pytest.approx(-np.expm1(-q) / q, rel=1e-12)
    q = 0.049
    _, _, A, C, D = O.drag_coefficients(q)
    em1, em2 = -np.expm1(-q), -np.expm1(-2 * q)
    assert A == pytest.approx(0.5 * em2, rel=2e-6)
    assert C == pytest.approx(em1 * em1 / q, rel=2e-6)
    # D from its Taylor series to high order (the closed form cancels badly here)
    Dser = q * sum(c * q ** k for k, c in enumerate([1 / 3, -1 / 4, 7 / 60, -1 / 24, 31 / 2520, -1 / 320,
                                                     127 / 181440]))
    assert D == pytest.approx(Dser, rel=2e-6)


@pytest.mark.parametrize("target", [(0.0, 0.0, 0.0), (0.1, -0.3, 0.05)])
def test_drag_on_a_uniform_state_is_the_exact_exponential_relaxation(target):
    """No gradients: fluxes cancel, so the step is the drag propagator alone. Velocity
    relaxes as v0 + (v - v0) e^{-q}; the energy lost equals the kinetic energy lost (the
    drag work), so internal energy, density and B are untouched."""
    n, q = 4, 0.37
    U = _uniform(n)
    s = O.OracleState((n, n, n), 1.0 / n, recon="plm_minmod")
    s.set_conserved(U)
    s.step_sources(1e-3, 1.0, 1.0, drag_impulse=q, target_velocity=target)
    v = U[1:4, 0] / U[0, 0]
    vn = s.U[1:4] / s.U[0]
    want = np.array(target) + (v - np.array(target)) * np.exp(-q)
    np.testing.assert_allclose(vn, np.broadcast_to(want[:, None], vn.shape), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(_eint(s.U), _eint(U), rtol=1e-13)
    np.testing.assert_allclose(s.U[[0, 5, 6, 7]], U[[0, 5, 6, 7]], rtol=1e-14)


def test_zero_impulse_drag_is_the_plain_step_and_zero_correction_radiation_is_the_drag_step():
    n = 8
    U = smooth_state(n, seed=3)
    a = O.OracleState((n, n, n), 1.0 / n, recon="plm")
    a.set_conserved(U)
    dt, ch = a.compute_dt(0.3)
    decay = 0.97
    b, c, d = a.copy(), a.copy(), a.copy()
    a.step_ref(dt, ch, decay)
    b.step_sources(dt, ch, decay, drag_impulse=0.0)
    scale = np.abs(a.U).max(axis=1, keepdims=True)
    assert (np.abs(b.U - a.U) / scale).max() < 1e-14
    c.step_sources(dt, ch, decay, drag_impulse=0.2)
    corr = np.zeros((9, n ** 3))
    corr[0] = U[8]  # psi enters through slot 0
    d.step_sources(dt, ch, decay, drag_impulse=0.2, correction=corr)
    np.testing.assert_array_equal(d.U, c.U)
    assert np.abs(c.U - a.U).max() > 1e-6


def test_radiation_velocity_correction_advects_along_its_own_axis():
    """A uniform velocity correction c_x on a density ripple at rest in x acts on the face
    states as an extra velocity: to first order the density moves as d rho = -c dt d_x rho,
    and a correction put on the y slot leaves this x-only ripple alone."""
    n, c, amp = 32, 0.05, 1e-3
    from pmfmhd import problems as P
    i = np.arange(n ** 3) % n
    x = 2 * np.pi * (i + 0.5) / n
    q = np.zeros((9, n ** 3))
    q[0] = 1 + amp * np.sin(x)
    q[4] = 1.0
    U = P.prim2cons(q, G)
    s = O.OracleState((n, n, n), 1.0 / n, recon="plm")
    s.set_conserved(U)
    dt = 0.2 / n
    for slot, moves in ((1, True), (2, False)):
        t = s.copy()
        corr = np.zeros((9, n ** 3))
        corr[slot] = c
        t.step_sources(dt, 1.0, 1.0, drag_impulse=0.0, correction=corr)
        drho = t.U[0] - U[0]
        want = -c * dt * amp * np.cos(x) * 2 * np.pi
        if moves:
            assert np.linalg.norm(drho - want) < 0.2 * np.linalg.norm(want)
        else:
            assert np.abs(drho).max() < 1e-3 * np.abs(want).max()


def test_tracking_carries_the_rounding_of_each_update():
    """Float32 with tracking: slot 4 accumulates the density increments (the result itself
    is unchanged), and the energy / B residual slots hold exactly the rounding error of
    each compensated update, so value + residual equals the float64 sum of the inputs."""
    n = 8
    U = smooth_state(n, seed=5).astype(np.float32)
    s = O.OracleState((n, n, n), 1.0 / n, dtype=np.float32, recon="plm")
    s.set_conserved(U)
    dt, ch = s.compute_dt(0.3)
    corr = np.zeros((9, n ** 3), dtype=np.float32)
    corr[0] = U[8]
    corr[5:9] = np.float32(1e-9)  # carried residuals from an earlier step
    plain = s.copy()
    plain.step_sources(dt, ch, 1.0, drag_impulse=0.0, correction=np.ascontiguousarray(corr.copy()))
    c0 = corr.copy()
    s.step_sources(dt, ch, 1.0, drag_impulse=0.0, correction=corr, track_density=True, track_magnetic=True,
                   track_energy=True)
    np.testing.assert_array_equal(s.U[0], plain.U[0])
    np.testing.assert_allclose(corr[4], (s.U[0].astype(np.float64) - U[0]), rtol=0, atol=2e-7)
    # exact Fast2Sum invariant for B: out + res_out = U0 + (du + res_in) with du recovered
    # from the untracked step (same fluxes): out_plain = U0 + du
    for v, slot in ((5, 5), (6, 6), (7, 7)):
        du = plain.U[v].astype(np.float64) - U[v]  # exact up to the plain rounding
        lhs = s.U[v].astype(np.float64) + corr[slot]
        rhs = U[v].astype(np.float64) + du + c0[slot]
        assert np.abs(lhs - rhs).max() < 4e-7 * np.abs(U[v]).max()
        assert np.abs(corr[slot]).max() < 1e-6  # a rounding error, not an increment


def test_admissibility_flag_marks_only_the_broken_cell():
    n = 8
    U = smooth_state(n, seed=2)
    c = 3 * n * n + 2 * n + 5
    U[4, c] = 0.5 * (U[1, c] ** 2 + U[2, c] ** 2 + U[3, c] ** 2) / U[0, c] + \
        0.5 * (U[5, c] ** 2 + U[6, c] ** 2 + U[7, c] ** 2) - 0.4  # negative internal energy
    a = O.OracleState((n, n, n), 1.0 / n, recon="plm_minmod", riemann="hll")
    a.set_conserved(U)
    b = a.copy()
    dt, ch = 1e-3, 2.0
    corr = np.zeros((9, n ** 3))
    corr[0] = U[8]
    a.step_sources(dt, ch, 1.0, drag_impulse=0.1, correction=corr.copy())
    b.step_sources(dt, ch, 1.0, drag_impulse=0.1, correction=corr.copy(), check=1)
    bad = b.U[0] < 0
    assert bad[c] and bad.sum() >= 1
    np.testing.assert_array_equal(b.U[:, ~bad], a.U[:, ~bad])
    np.testing.assert_allclose(-b.U[0, bad], np.maximum(np.abs(a.U[0, bad]), 1e-4))


def test_global_fallback_equals_masked_fallback_with_every_cell_flagged():
    n = 8
    U = smooth_state(n, seed=4)
    a = O.OracleState((n, n, n), 1.0 / n, recon="plm", riemann="hlld")
    a.set_conserved(U)
    b = a.copy()
    dt, ch = a.compute_dt(0.3)
    corr = np.zeros((9, n ** 3))
    corr[0] = U[8]
    ca, cb = corr.copy(), corr.copy()
    cb[4] = 1.0
    a.step_sources(dt, ch, 0.9, drag_impulse=0.05, correction=ca, check=3, riemann="llf")
    b.step_sources(dt, ch, 0.9, drag_impulse=0.05, correction=cb, check=2)
    np.testing.assert_array_equal(a.U, b.U)
    # and a flagged uniform block stays first order: zero slopes give the cell-mean faces
    c = O.OracleState((n, n, n), 1.0 / n, recon="plm")
    c.set_conserved(_uniform(n))
    cc = np.zeros((9, n ** 3))
    cc[4] = 1.0
    c.step_sources(dt, ch, 1.0, drag_impulse=0.0, correction=cc, check=2)
    np.testing.assert_allclose(c.U, _uniform(n), rtol=1e-14, atol=1e-15)


def test_drag_propagates_the_face_velocity_to_the_half_step():
    """A density ripple advected by a uniform velocity u: with a drag impulse q the faces
    carry the half-step velocity u e^{-q/2} (the propagated face prediction), so the
    density change is the plain one scaled by e^{-q/2} (0.61 at q = 1; phi1(q/2) = 0.79 or
    no propagation = 1 would be far off)."""
    n, u, amp, q = 32, 0.5, 1e-3, 1.0
    from pmfmhd import problems as P
    i = np.arange(n ** 3) % n
    x = 2 * np.pi * (i + 0.5) / n
    qp = np.zeros((9, n ** 3))
    qp[0] = 1 + amp * np.sin(x)
    qp[1] = u
    qp[4] = 1.0
    U = P.prim2cons(qp, G)
    s = O.OracleState((n, n, n), 1.0 / n, recon="plm")
    s.set_conserved(U)
    dt = 0.1 / n
    a, b = s.copy(), s.copy()
    a.step_sources(dt, 1.0, 1.0, drag_impulse=0.0)
    b.step_sources(dt, 1.0, 1.0, drag_impulse=q)
    ratio = np.linalg.norm(b.U[0] - U[0]) / np.linalg.norm(a.U[0] - U[0])
    assert ratio == pytest.approx(np.exp(-q / 2), rel=0.03)


# ---------------------------------------------------------------------------
# GPU kernels against the oracle (float64: 1e-12; float32: 1e-5)

def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.maximum(np.abs(b).max(axis=-1, keepdims=True), 1e-30)
    return float((np.abs(a - b) / scale).max())


def _corr(n, U, seed=0, dtype=np.float64):
    rng = np.random.default_rng(seed)
    c = np.zeros((9, n ** 3))
    c[0] = U[8]
    c[1:4] = rng.normal(0, 0.02, (3, n ** 3))
    c[4] = rng.normal(0, 1e-3, n ** 3)
    c[5:9] = rng.normal(0, 1e-7, (4, n ** 3))
    return np.ascontiguousarray(c.astype(dtype))


def _pair(n, U, dtype, recon, riemann, bcs=("periodic",) * 3, solver="glm"):
    import pmfmhd as M
    s = M.allocate_state((n, n, n), 1.0 / n, dtype=dtype, recon=recon, riemann=riemann, bcs=bcs, mhd_solver=solver)
    o = O.OracleState((n, n, n), 1.0 / n, dtype=dtype, recon=recon, riemann=riemann, bcs=bcs, mhd_solver=solver)
    s.set_conserved(U)
    o.set_conserved(U)
    return s, o


@pytest.mark.gpu
@pytest.mark.parametrize("integ", ["ref", "cube"])
@pytest.mark.parametrize("recon,riemann", [("plm", "hlld"), ("ppm", "hll"), ("plm_minmod", "llf")])
def test_drag_step_matches_oracle_float64(integ, recon, riemann):
    import pmfmhd as M
    n = 16
    U = smooth_state(n, seed=7)
    s, o = _pair(n, U, np.float64, recon, riemann)
    dt, ch = o.compute_dt(0.3)
    for _ in range(3):
        M.step_drag(s, dt, ch, 0.95, 0.3, integrator=integ)
        o.step_sources(dt, ch, 0.95, drag_impulse=0.3)
    assert _rel(s.conserved(), o.U) < 1e-12


@pytest.mark.gpu
def test_ref_drag_with_a_moving_target_and_boundaries_matches_oracle():
    import pmfmhd as M
    n = 12
    U = smooth_state(n, seed=8)
    s, o = _pair(n, U, np.float64, "plm", "hlld", bcs=("outflow", "reflecting", "periodic"))
    dt, ch = o.compute_dt(0.3)
    v0 = (0.1, -0.2, 0.05)
    for _ in range(2):
        M.step_drag(s, dt, ch, 0.97, 0.8, target_velocity=v0, integrator="ref")
        o.step_sources(dt, ch, 0.97, drag_impulse=0.8, target_velocity=v0)
    assert _rel(s.conserved(), o.U) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("bcs", [("outflow", "reflecting", "periodic"), ("reflecting", "reflecting", "outflow")])
def test_cube_coupled_steps_on_bounded_grids_match_oracle(bcs):
    """step_cube_drag! / step_cube_radiation! on outflow / reflecting grids
    (integrator_cube.jl:478-540 with per = false): ghost cells from the boundary map, the
    radiation corrections and masks at the periodic image of each cell
    (integrator_ref.jl:50-53)."""
    import pmfmhd as M
    n = 16
    U = smooth_state(n, seed=12)
    s, o = _pair(n, U, np.float64, "ppm", "hlld", bcs=bcs)
    dt, ch = o.compute_dt(0.3)
    M.step_drag(s, dt, ch, 0.95, 0.3, integrator="cube")
    o.step_sources(dt, ch, 0.95, drag_impulse=0.3)
    assert _rel(s.conserved(), o.U) < 1e-12
    c = _corr(n, o.U, seed=5)
    c[4] = (np.random.default_rng(6).random(n ** 3) < 0.1).astype(np.float64)
    buf = M.correction_buffer(s, c)
    M.step_radiation(s, dt, ch, 0.9, 0.2, buf, fallback=True, track_density=False, integrator="cube")
    o.step_sources(dt, ch, 0.9, drag_impulse=0.2, correction=c, track_density=False, track_magnetic=True,
                   track_energy=True, check=2)
    assert _rel(s.conserved(), o.U) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("integ", ["ref", "cube"])
def test_radiation_step_with_tracking_matches_oracle_float64(integ):
    import pmfmhd as M
    n = 16
    U = smooth_state(n, seed=9)
    s, o = _pair(n, U, np.float64, "plm", "hlld")
    c = _corr(n, U, seed=1)
    buf = M.correction_buffer(s, c)
    dt, ch = o.compute_dt(0.3)
    for _ in range(2):
        M.step_radiation(s, dt, ch, 0.9, 0.4, buf, integrator=integ)
        o.step_sources(dt, ch, 0.9, drag_impulse=0.4, correction=c, track_density=True, track_magnetic=True,
                       track_energy=True)
        c[0] = o.U[8]  # the radiation loop reloads psi into slot 0 before each step (runner.jl)
        M.load_psi(s, buf)
    assert _rel(s.conserved(), o.U) < 1e-12
    got = buf.download(np.float64, (9, n ** 3))
    assert _rel(got[1:4], c[1:4]) == 0.0
    assert _rel(got[4], c[4]) < 1e-12
    # residual slots hold rounding errors: compare in absolute terms against the field scale
    for v in (5, 6, 7, 8):
        assert np.abs(got[v] - c[v]).max() < 1e-13 * max(1.0, np.abs(o.U[v if v < 8 else 4]).max())


@pytest.mark.gpu
def test_ct_radiation_step_matches_oracle_float64():
    import pmfmhd as M
    from test_oracle import ct_test_hydro, ct_test_potential
    n = 12
    U = ct_test_hydro(n, G)
    s, o = _pair(n, U, np.float64, "plm", "hll", solver="ct")
    Ax, Ay, Az = ct_test_potential(n)
    M.ct_faces_from_potential(s, Ax, Ay, Az)
    o.ct_faces_from_potential(Ax, Ay, Az)
    c = _corr(n, o.U, seed=2)
    buf = M.correction_buffer(s, c)
    dt, _ = o.compute_dt(0.3)
    for _ in range(2):
        M.step_radiation(s, dt, 0.0, 1.0, 0.2, buf, integrator="ref")
        o.step_sources(dt, 0.0, 1.0, drag_impulse=0.2, correction=c, track_density=True, track_magnetic=True,
                       track_energy=True)
    assert _rel(s.conserved(), o.U) < 1e-12
    assert _rel(s.ct_slot(0), o.ct[0]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("recon,riemann", [("plm", "hlld"), ("ppm", "hll"), ("plm_minmod", "llf")])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ct_coupled_steps_fused_equal_per_cell_kernel(recon, riemann, dtype):
    """step_ref_drag! / step_ref_radiation! on CT states run on the fused CT kernel (every
    face once) and on the per-cell kernel (pmf_set_ct_kernel): the same cells, faces,
    workspace slots (face fluxes, cell EMFs from the source-modified velocity) and
    correction-buffer slots, to round-off; float64 also against the oracle."""
    import pmfmhd as M
    from test_oracle import ct_test_hydro, ct_test_potential
    n = 16
    U = ct_test_hydro(n, G)
    A = ct_test_potential(n)
    runs = []
    for per_cell in (False, True):
        s = M.allocate_state((n,) * 3, 1.0 / n, dtype=dtype, recon=recon, riemann=riemann, mhd_solver="ct")
        s.set_conserved(U)
        M.ct_faces_from_potential(s, *A)
        M.set_ct_kernel(s, per_cell)
        c = _corr(n, U, seed=7, dtype=dtype)
        c[0] = 0.0
        buf = M.correction_buffer(s, c)
        dt, _ = M.compute_dt(s, 0.3)
        M.step_drag(s, float(dt), 0.0, 1.0, 0.3, integrator="ref")
        for _ in range(2):
            M.step_radiation(s, float(dt), 0.0, 1.0, 0.2, buf, integrator="ref")
        runs.append((s.conserved(), [s.ct_slot(k) for k in range(15)],
                     buf.download(np.float64 if dtype == np.float64 else np.float32, (9, n ** 3)), float(dt)))
    tol = 1e-12 if dtype == np.float64 else 2e-6
    (Uf, Sf, Bf, dtf), (Ur, Sr, Br, _) = runs
    assert _rel(Uf, Ur) < tol
    for k in range(15):
        assert _rel(Sf[k][None], Sr[k][None]) < tol, k
    # slot 4 accumulates density increments, differences of face fluxes: in float32 the
    # two kernels' roundings of those fluxes show there at a few 1e-6 of its largest value
    assert _rel(Bf[1:5], Br[1:5]) < (tol if dtype == np.float64 else 2e-5)
    if dtype == np.float64:
        o = O.OracleState((n,) * 3, 1.0 / n, recon=recon, riemann=riemann, mhd_solver="ct")
        o.set_conserved(U)
        o.ct_faces_from_potential(*A)
        c = _corr(n, U, seed=7)
        c[0] = 0.0
        o.step_sources(dtf, 0.0, 1.0, drag_impulse=0.3)
        for _ in range(2):
            o.step_sources(dtf, 0.0, 1.0, drag_impulse=0.2, correction=c, track_density=True, track_magnetic=True,
                           track_energy=True)
        assert _rel(Uf, o.U) < 1e-12
        for d in range(3):
            assert _rel(Sf[d][None], o.ct[d][None]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["check", "fallback", "global", "robust"])
def test_cube_radiation_checks_and_fallbacks_match_oracle(mode):
    import pmfmhd as M
    n = 16
    U = smooth_state(n, seed=11)
    cbad = 5 * n * n + 3 * n + 9
    U[4, cbad] = 0.5 * (U[1, cbad] ** 2 + U[2, cbad] ** 2 + U[3, cbad] ** 2) / U[0, cbad] + \
        0.5 * (U[5, cbad] ** 2 + U[6, cbad] ** 2 + U[7, cbad] ** 2) - 0.3
    s, o = _pair(n, U, np.float64, "plm", "hlld")
    c = _corr(n, U, seed=3)
    if mode == "fallback":
        c[4] = (np.random.default_rng(4).random(n ** 3) < 0.1).astype(np.float64)
    buf = M.correction_buffer(s, c)
    dt, ch = 2e-3, 2.0
    kw = dict(check_admissibility=mode == "check", fallback=mode == "fallback", global_fallback=mode == "global",
              robust_fallback=mode == "robust", track_density=mode != "fallback")
    M.step_radiation(s, dt, ch, 0.9, 0.1, buf, integrator="cube", **kw)
    o.step_sources(dt, ch, 0.9, drag_impulse=0.1, correction=c, track_density=mode != "fallback",
                   track_magnetic=True, track_energy=True,
                   check={"check": 1, "fallback": 2, "global": 3, "robust": 0}[mode],
                   recon="plm_minmod" if mode == "robust" else None,
                   riemann={"global": "llf", "robust": "hll"}.get(mode))
    got = s.conserved()
    assert _rel(got, o.U) < 1e-12
    if mode == "check":
        assert got[0, cbad] < 0 and (got[0] < 0).sum() == (o.U[0] < 0).sum()


@pytest.mark.gpu
@pytest.mark.parametrize("integ", ["ref", "cube"])
def test_radiation_float32_against_float32_oracle(integ):
    import pmfmhd as M
    from pmfmhd import problems as P
    n = 32
    U = P.config_state("bstatic_32", seed=4).astype(np.float32)
    s, o = _pair(n, U, np.float32, "plm_minmod", "hlld")
    c = _corr(n, U, seed=5, dtype=np.float32)
    buf = M.correction_buffer(s, c)
    dt, ch = o.compute_dt(0.2)
    M.step_radiation(s, dt, ch, 0.99, 0.05, buf, integrator=integ)
    o.step_sources(dt, ch, 0.99, drag_impulse=0.05, correction=c, track_density=True, track_magnetic=True,
                   track_energy=True)
    got = s.conserved().astype(np.float64)
    ref = o.U.astype(np.float64)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5


@pytest.mark.gpu
def test_drag_ref_and_cube_agree_float32():
    """The reference's own check (test_cube_cpu.jl:92-101): the drag step is the same on
    both integrators."""
    import pmfmhd as M
    n = 16
    U = smooth_state(n, seed=12).astype(np.float32)
    a, _ = _pair(n, U, np.float32, "ppm", "hlld")
    b, _ = _pair(n, U, np.float32, "ppm", "hlld")
    dt, ch = M.compute_dt(a, 0.08)
    for _ in range(3):
        M.step_drag(a, dt, ch, 0.99, 0.05, integrator="ref")
        M.step_drag(b, dt, ch, 0.99, 0.05, integrator="cube")
    ua, ub = a.conserved().astype(np.float64), b.conserved().astype(np.float64)
    assert np.abs(ua - ub).max() / np.abs(ua).max() < 2e-6
    mass0 = U[0].astype(np.float64).sum()
    assert abs(ub[0].sum() / mass0 - 1) < 3e-6


@pytest.mark.gpu
def test_coupled_step_error_texts():
    import pmfmhd as M
    n = 16
    U = smooth_state(n, seed=1)
    s, _ = _pair(n, U, np.float32, "plm", "hlld")
    buf = M.correction_buffer(s)
    with pytest.raises(M.MHDError, match="specialized for Compton drag in the CMB frame"):
        M.step_drag(s, 1e-3, 1.0, 1.0, 0.1, target_velocity=(0.1, 0, 0), integrator="cube")
    with pytest.raises(M.MHDError, match="options of the :cube integrator"):
        M.step_radiation(s, 1e-3, 1.0, 1.0, 0.1, buf, check_admissibility=True, integrator="ref")
    with pytest.raises(M.MHDError, match="density tracking writes"):
        M.step_radiation(s, 1e-3, 1.0, 1.0, 0.1, buf, fallback=True, integrator="cube")
    with pytest.raises(M.MHDError, match="needs the GLM cleaning speed"):
        M.step_drag(s, 1e-3, 0.0, 1.0, 0.1, integrator="ref")
    c, _ = _pair(n, U, np.float32, "plm", "hll", solver="ct")
    with pytest.raises(M.MHDError, match="step_cube_drag!: CT is implemented on the :ref integrator only"):
        M.step_drag(c, 1e-3, 0.0, 1.0, 0.1, integrator="cube")


def _device_fields(M, arrays):
    bufs = []
    for a in arrays:
        a = np.ascontiguousarray(a)
        b = M.DeviceBuffer(a.nbytes)
        b.upload(a)
        bufs.append(b)
    return bufs


_SEAM_CASES = {
    # name: (solver, integrator, radiation, flags)
    "drag-ref": ("glm", "ref", False, 0),
    "drag-cube": ("glm", "cube", False, 0),
    "radiation-ref-tracked": ("glm", "ref", True, 7),
    "radiation-cube-check": ("glm", "cube", True, 7 | 8),
    "radiation-cube-fallback": ("glm", "cube", True, 6 | 16),
    "radiation-cube-robust": ("glm", "cube", True, 1 | 64),
    "ct-drag": ("ct", "ref", False, 0),
    "ct-radiation-tracked": ("ct", "ref", True, 7),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(_SEAM_CASES))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sources_into_caller_arrays_equal_the_state_steps(case, dtype):
    """The _cube_drag_launch! / _cube_radiation_launch! seam (pmf_step_sources_into): nine
    caller arrays stepped through a field-less workspace into nine others, with the caller's
    correction buffer and, for CT, its 21-slot packed buffer, give the same bits as
    pmf_step_drag / pmf_step_radiation on a state holding the same fields: cells, correction
    slots and CT slots alike."""
    import pmfmhd as M
    from test_oracle import ct_test_hydro, ct_test_potential
    solver, integ, radiation, flags = _SEAM_CASES[case]
    n = 16
    ct = solver == "ct"
    riemann = "hll" if ct else "hlld"
    U = ct_test_hydro(n, G) if ct else smooth_state(n, seed=13)
    s, _ = _pair(n, U, dtype, "plm", riemann, solver=solver)
    if ct:
        M.ct_faces_from_potential(s, *ct_test_potential(n))
    dt, smax = M.compute_dt(s, 0.3)
    ch = 0.0 if ct else float(smax)
    decay = 1.0 if ct else 0.93
    c = _corr(n, s.conserved(), seed=17, dtype=dtype)
    if flags & 16:
        c[4] = (np.random.default_rng(18).random(n ** 3) < 0.1).astype(dtype)
    U0 = s.conserved()
    slots0 = [s.ct_slot(q) for q in range(21)] if ct else None
    # the seam's inputs: copies of the state before the step
    ins = _device_fields(M, list(U0))
    outs = _device_fields(M, [np.zeros_like(U0[0])] * 9)
    ctbuf = _device_fields(M, [np.concatenate(slots0)])[0] if ct else None
    corr_seam = M.correction_buffer(s, c) if radiation else None
    corr_state = M.correction_buffer(s, c) if radiation else None
    ws = M.allocate_state((n,) * 3, 1.0 / n, dtype=dtype, recon="plm", riemann=riemann, mhd_solver=solver,
                          workspace=True)
    M.step_sources_into(ws, [b.ptr for b in ins], [b.ptr for b in outs], float(dt), ch, decay, 0.3,
                        correction=corr_seam, flags=flags, integrator=integ,
                        ct_ptr=None if ctbuf is None else ctbuf.ptr)
    ws.synchronize()
    if radiation:
        kw = dict(track_density=bool(flags & 1), track_magnetic=bool(flags & 2), track_energy=bool(flags & 4),
                  check_admissibility=bool(flags & 8), fallback=bool(flags & 16), robust_fallback=bool(flags & 64))
        M.step_radiation(s, float(dt), ch, decay, 0.3, corr_state, integrator=integ, **kw)
    else:
        M.step_drag(s, float(dt), ch, decay, 0.3, integrator=integ)
    got = np.stack([b.download(dtype, n ** 3) for b in outs])
    assert np.array_equal(got, s.conserved())
    if radiation:
        assert np.array_equal(corr_seam.download(dtype, (9, n ** 3)), corr_state.download(dtype, (9, n ** 3)))
    if ct:
        slots = ctbuf.download(dtype, (21, n ** 3))
        for q in range(21):  # faces, mass fluxes, face fluxes, cell and edge EMFs, remainders
            assert np.array_equal(slots[q], s.ct_slot(q)), q


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_sources_into_other_arrays_leaves_the_kept_cfl_speed_alone(dtype):
    """A coupled step of other arrays through a state must not leave their CFL speed
    standing in for the state's own (the fused epilogue writes it to the state's device
    scalar)."""
    import pmfmhd as M
    n = 16
    s, _ = _pair(n, smooth_state(n, seed=3), dtype, "plm", "hlld")
    dt, sm = M.compute_dt(s, 0.3)
    M.step(s, float(dt), ch=float(sm), integrator="cube")
    _, kept = M.compute_dt(s, 0.3)
    other = 2.0 * smooth_state(n, seed=4)
    ins = _device_fields(M, list(other.astype(dtype)))
    outs = _device_fields(M, [np.zeros(n ** 3, dtype)] * 9)
    M.step_sources_into(s, [b.ptr for b in ins], [b.ptr for b in outs], 1e-3, 3.0, 1.0, 0.2, integrator="cube")
    _, again = M.compute_dt(s, 0.3)
    assert again == kept
