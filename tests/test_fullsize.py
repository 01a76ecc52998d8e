"""Size-independent properties at the benchmark size (256^3 float32, the
plain-turbulence workload): exact translation equivariance of the fused step,
conservation of the periodic totals, agreement of the fused kernels with the per-cell
(:ref) kernels for GLM and CT states, the coupled steps, and species transport."""
import numpy as np
import pytest

import pmfmhd as M
from pmfmhd import problems as P

pytestmark = pytest.mark.gpu
N = 256


@pytest.fixture(scope="module")
def state256():
    return P.config_state("turb_256", seed=1).astype(np.float32)


def _state(U, **kw):
    s = M.allocate_state((N, N, N), 1.0 / N, dtype=np.float32, recon="plm_minmod", **kw)
    s.set_conserved(U)
    return s


def test_fused_step_is_translation_equivariant_bitwise(state256):
    """Every face is solved by the same arithmetic wherever it falls (tile interior, tile
    ring, z-chunk seam), so a periodic shift of the input shifts the output exactly."""
    U = state256
    a = _state(U)
    dt, sm = M.compute_dt(a, 0.2)
    M.step(a, float(dt), ch=float(sm), integrator="cube")
    shift = (3, 7, 5)  # cells along z, y, x: not multiples of the 12 x 16 tile or the chunk
    Us = np.roll(U.reshape(9, N, N, N), shift, axis=(1, 2, 3)).reshape(9, -1)
    b = _state(np.ascontiguousarray(Us))
    M.step(b, float(dt), ch=float(sm), integrator="cube")
    ra = np.roll(a.conserved().reshape(9, N, N, N), shift, axis=(1, 2, 3)).reshape(9, -1)
    assert np.array_equal(ra, b.conserved())


def test_fused_steps_conserve_the_periodic_totals(state256):
    s = _state(state256)
    t0 = M.conserved_totals(s)
    for _ in range(5):
        dt, sm = M.compute_dt(s, 0.2)
        M.step(s, float(dt), ch=float(sm), integrator="cube")
    t1 = M.conserved_totals(s)
    for k in ("mass", "energy"):
        assert abs(t1[k] - t0[k]) <= 1e-9 * abs(t0[k]), k
    scale = np.sqrt(t0["energy"])  # momentum and field totals start near zero
    for k in ("momx", "momy", "momz", "Bx", "By", "Bz"):
        assert abs(t1[k] - t0[k]) <= 1e-9 * scale, k
    assert all(np.isfinite(a).all() for a in s.fields_to_host())


def test_fused_and_per_cell_glm_steps_agree(state256):
    a, b = _state(state256), _state(state256)
    dt, sm = M.compute_dt(a, 0.2)
    M.step(a, float(dt), ch=float(sm), integrator="cube")
    M.step(b, float(dt), ch=float(sm), integrator="ref")
    x, y = a.conserved().astype(np.float64), b.conserved().astype(np.float64)
    for v in range(8):
        assert np.max(np.abs(x[v] - y[v])) <= 2e-6 * np.abs(y[v]).max(), v


def test_fused_and_per_cell_ct_steps_agree(state256):
    a, b = _state(state256, mhd_solver="ct"), _state(state256, mhd_solver="ct")
    M.set_ct_kernel(b, True)
    for s in (a, b):
        M.ct_faces_from_cells(s)
    dt, _ = M.compute_dt(a, 0.2)
    for s in (a, b):
        M.step(s, float(dt))
    x, y = a.conserved().astype(np.float64), b.conserved().astype(np.float64)
    for v in range(8):
        assert np.max(np.abs(x[v] - y[v])) <= 2e-6 * np.abs(y[v]).max(), v
    assert M.max_divb_face(a) / N < 1e-5


def test_coupled_steps_at_full_size(state256):
    """Drag / radiation-coupled fused steps at 256^3: radiation with a zero correction
    buffer and tracking off is the drag step bit for bit, and the drag step keeps the
    mass total while removing energy."""
    a, b = _state(state256), _state(state256)
    dt, sm = M.compute_dt(a, 0.2)
    dt, sm = float(dt), float(sm)
    decay = 0.99
    t0 = M.conserved_totals(a)
    M.step_drag(a, dt, sm, decay, 0.2, integrator="cube")
    buf = M.correction_buffer(b)
    M.load_psi(b, buf)
    M.step_radiation(b, dt, sm, decay, 0.2, buf, track_density=False, track_magnetic=False, track_energy=False,
                     integrator="cube")
    assert np.array_equal(a.conserved(), b.conserved())
    t1 = M.conserved_totals(a)
    assert abs(t1["mass"] - t0["mass"]) <= 1e-9 * t0["mass"]
    assert t1["energy"] < t0["energy"]


def test_species_transport_at_full_size(state256):
    """Passive species on the 256^3 flow (fractions in, densities out): a uniform
    composition comes back as exactly rho * fraction, and random coupled fractions stay
    within [0, maximum_total] of the density."""
    s = _state(state256)
    nc = N ** 3
    rho = np.maximum(state256[0], np.finfo(np.float32).eps)
    f0 = np.float32(0.25)
    b1, t1 = M.DeviceBuffer(4 * nc), M.DeviceBuffer(4 * nc)
    b1.upload(np.full(nc, f0, np.float32))
    M.advect_species(s, b1, t1, 2e-3, maximum_total=1.0)
    assert np.array_equal(b1.download(np.float32, nc), rho * f0)
    rng = np.random.default_rng(7)
    b2, t2 = M.DeviceBuffer(4 * nc), M.DeviceBuffer(4 * nc)
    b1.upload((state256[0] * 0.6 * rng.random(nc)).astype(np.float32))
    b2.upload((state256[0] * 0.6 * rng.random(nc)).astype(np.float32))
    M.species_to_fractions(s, b1, b2, maximum_total=0.9)
    M.advect_species(s, b1, t1, 2e-3, species2=b2, scratch2=t2, maximum_total=0.9)
    d1, d2 = b1.download(np.float32, nc), b2.download(np.float32, nc)
    assert d1.min() >= 0 and d2.min() >= 0
    assert ((d1.astype(np.float64) + d2) / rho).max() <= 0.9 * (1 + 1e-6)


@pytest.mark.timeout(900)
def test_mach15_512_fused_equals_per_cell_and_keeps_the_ledger():
    """512^3, the bandwidth-roofline configuration (Mach-1.5 solenoidal velocity, lognormal
    rho sigma 0.3, |m| <= 64, rms|B| 1): one fused step equals the per-cell kernel's to
    2e-6 of each field's maximum; five fused steps keep every field finite, the density
    positive, and the float64 ledger's mass and energy totals to 1e-9."""
    n = 512
    print("mach15 512^3: building the initial state", flush=True)
    U = P.config_state("mach15_512", seed=1).astype(np.float32)
    a = M.allocate_state((n,) * 3, 1.0 / n, dtype=np.float32, recon="plm_minmod")
    b = M.allocate_state((n,) * 3, 1.0 / n, dtype=np.float32, recon="plm_minmod")
    a.set_conserved(U)
    b.set_conserved(U)
    del U
    t0 = M.conserved_totals(a)
    dt, sm = M.compute_dt(a, 0.2)
    M.step(a, float(dt), ch=float(sm), integrator="cube")
    M.step(b, float(dt), ch=float(sm), integrator="ref")
    for v in range(9):
        x, y = a.download(v).astype(np.float64), b.download(v).astype(np.float64)
        ref = np.abs(y).max() if v < 8 else np.abs(b.download(5)).max()
        assert np.max(np.abs(x - y)) <= 2e-6 * ref, v
    del b, x, y
    for _ in range(4):
        dt, sm = M.compute_dt(a, 0.2)
        M.step(a, float(dt), ch=float(sm), integrator="cube")
    t1 = M.conserved_totals(a)
    for k in ("mass", "energy"):
        assert abs(t1[k] - t0[k]) <= 1e-9 * abs(t0[k]), k
    rho = a.download(0)
    assert np.isfinite(rho).all() and rho.min() > 0
    for v in range(1, 9):
        assert np.isfinite(a.download(v)).all(), v


def test_surrogate_step_at_full_size_without_gated_faces_is_the_plain_step():
    """256^3 on the surrogate workload's inputs (|m| <= 8, rms|B| 0.3, v = 0): with the gate
    closed (thr = 0) no face takes the network and the surrogate kernel's step equals the
    plain fused step to float32 rounding, with the same CFL speed."""
    from pmfmhd.surrogate import Surrogate, step_surrogate
    U = P.config_state("bstatic_256", seed=1).astype(np.float32)
    a, b = _state(U), _state(U)
    net = Surrogate(b, seed=1)
    dt, smax = M.compute_dt(a, 0.2)
    M.step(a, float(dt), ch=float(smax), integrator="cube")
    assert step_surrogate(b, net, float(dt), float(smax), thr=0.0, count=True) == 0
    # the momenta start at zero, so they are measured on the momentum scale of the state,
    # sqrt(max rho * max E), and psi on the field's
    mscale = np.sqrt(np.abs(a.download(0)).max() * np.abs(a.download(4)).max())
    for v in range(9):
        x, y = a.download(v).astype(np.float64), b.download(v).astype(np.float64)
        ref = np.abs(x).max() if v in (0, 4, 5, 6, 7) else (mscale if v < 4 else np.abs(a.download(5)).max())
        assert np.max(np.abs(x - y)) <= 2e-6 * ref, v
    assert M.max_wavespeed(a) == pytest.approx(M.max_wavespeed(b), rel=1e-6)
