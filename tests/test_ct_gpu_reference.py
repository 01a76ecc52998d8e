"""The reference's constrained-transport checks run by the HIP library
(test/inherited/test_ct.jl): second-order convergence of the CT step on an Alfvén wave
and an entropy wave (367-385), and the compensated face updates under a strong guide
field (470-564). The grids have nx a multiple of 16, so the plain CT steps run on the
fused CT kernel; the radiation-coupled CT steps on the per-cell kernel."""
import numpy as np
import pytest

import pmfmhd as M
from pmfmhd import problems as P

pytestmark = pytest.mark.gpu
G = 5.0 / 3.0


def _ct_alfven_L1(dtype, N, amp=0.1, p0=0.1, tfinal=0.5, cfl=0.4):
    s = M.allocate_state((N, 4, 4), 1.0 / N, dtype=dtype, mhd_solver="ct")
    s.set_conserved(P.alfven_wave((N, 4, 4), 1.0 / N, amp=amp, B0=1.0, p0=p0))
    M.ct_faces_from_cells(s)
    t, n = M.evolve(s, tfinal, cfl=cfl)
    By = s.download(6)[:N].astype(np.float64)
    x = (np.arange(N) + 0.5) / N
    return np.mean(np.abs(By - P.alfven_By_exact(x, t, 1.0, amp=amp))), n


def _ct_entropy_L1(dtype, N, tfinal=1.0, cfl=0.4):
    x = (np.arange(N) + 0.5) / N
    rho = np.tile(1 + 0.2 * np.sin(2 * np.pi * x), 16)
    U = np.zeros((9, 16 * N))
    U[0], U[1], U[2] = rho, rho * 1.0, rho * 0.3
    U[4] = 1 / (G - 1) + 0.5 * rho * (1.0 + 0.09) + 0.5 * (0.25 + 0.09)
    U[5], U[6] = 0.5, 0.3
    s = M.allocate_state((N, 4, 4), 1.0 / N, dtype=dtype, mhd_solver="ct")
    s.set_conserved(U)
    M.ct_faces_from_cells(s)
    t, n = M.evolve(s, tfinal, cfl=cfl)
    r = s.download(0)[:N].astype(np.float64)
    return np.mean(np.abs(r - (1 + 0.2 * np.sin(2 * np.pi * np.mod(x - t, 1.0))))), n


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ct_second_order_convergence(dtype):
    for f, Ns in ((_ct_alfven_L1, (16, 32, 64)), (_ct_entropy_L1, (32, 64, 128))):
        L1 = [f(dtype, N)[0] for N in Ns]
        assert all(np.isfinite(L1))
        orders = [np.log2(L1[k - 1] / L1[k]) for k in (1, 2)]
        assert min(orders) >= 1.8, (f.__name__, L1, orders)


def _guide_problem(N, P0=1000.0, B0=(35.0, 22.0, 28.0), amp=2e-4, vamp=1e-4):
    """test_ct.jl:478-505 (float32 inputs)."""
    f32 = np.float32
    k, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    x, y, z = i / N, j / N, k / N
    Ax, Ay, Az = (np.zeros((N, N, N), f32) for _ in range(3))
    vx, vy, vz = (np.zeros((N, N, N), f32) for _ in range(3))
    rho = np.ones((N, N, N), f32)
    modes = ((1, 2, -1, 0.3), (2, -1, 1, 1.1), (-1, 1, 2, 2.0), (3, 1, -2, 0.7), (1, -3, 2, 1.9))
    for m, (a, b, c, ph) in enumerate(modes, start=1):
        arg = 2 * np.pi * (a * x + b * y + c * z) + ph
        Ax += (amp * np.sin(arg + m) / (2 * np.pi * (abs(b) + abs(c)))).astype(f32)
        Ay += (amp * np.cos(arg) / (2 * np.pi * (abs(a) + abs(c)))).astype(f32)
        Az += (amp * np.sin(arg + 2 * m) / (2 * np.pi * (abs(a) + abs(b)))).astype(f32)
        vx += (vamp * np.sin(arg + 0.5 * m)).astype(f32)
        vy += (vamp * np.cos(arg + m)).astype(f32)
        vz += (vamp * np.sin(arg + 1.5 * m)).astype(f32)
        rho += (1e-3 * np.cos(arg + 2 * m)).astype(f32)
    hyd = [rho, rho * vx, rho * vy, rho * vz,
           f32(P0 / (G - 1)) + f32(0.5) * rho * (vx * vx + vy * vy + vz * vz)]
    g = M.allocate_state((N, N, N), 1.0 / N, dtype=np.float32, mhd_solver="ct")
    M.ct_faces_from_potential(g, Ax.reshape(-1), Ay.reshape(-1), Az.reshape(-1))
    faces = [g.ct_slot(d) + f32(B0[d]) for d in range(3)]
    return faces, [h.reshape(-1) for h in hyd]


def _guide_state(dtype, N, faces, hyd):
    s = M.allocate_state((N, N, N), 1.0 / N, dtype=dtype, riemann="hlld", mhd_solver="ct")
    U = np.zeros((9, N ** 3), dtype)
    for v in range(5):
        U[v] = hyd[v]
    s.set_conserved(U)
    M.ct_install_faces(s, *[f.astype(dtype) for f in faces])
    return s


def test_ct_compensated_faces_under_a_strong_guide_field():
    """Float32 faces of ~35 (ulp 3.8e-6) with per-step increments near half an ulp:
    plain summation lets the face divergence random-walk and the perturbation drift from
    float64; the compensated update (radiation-coupled CT step with track_magnetic)
    keeps both at round-off."""
    N, nsteps = 16, 2000
    faces, hyd = _guide_problem(N)
    sp = _guide_state(np.float32, N, faces, hyd)   # plain step (fused CT kernel)
    sn = _guide_state(np.float32, N, faces, hyd)   # radiation kernel, track_magnetic off
    sc = _guide_state(np.float32, N, faces, hyd)   # radiation kernel, track_magnetic on
    s64 = _guide_state(np.float64, N, faces, hyd)
    dt, _ = M.compute_dt(sp, 0.3)
    dt = float(np.float32(dt))
    bn = M.correction_buffer(sn)
    bc = M.correction_buffer(sc)

    def brms64(s):
        h = [s.download(v).astype(np.float64) for v in (5, 6, 7)]
        return np.sqrt(sum(np.sum(a * a) for a in h) / h[0].size)

    def divn(s):
        return M.max_divb_face(s) * (1.0 / N) / brms64(s)

    b0 = [s64.download(5 + d).astype(np.float64) for d in range(3)]
    div0 = divn(sp)
    for _ in range(nsteps):
        M.step(sp, dt)
        M.step_radiation(sn, dt, 0.0, 1.0, 0.0, bn, track_density=False, track_magnetic=False, track_energy=False)
        M.step_radiation(sc, dt, 0.0, 1.0, 0.0, bc, track_density=False, track_magnetic=True, track_energy=False)
        M.step(s64, dt)
    B64 = [s64.download(5 + d).astype(np.float64) for d in range(3)]
    pert = max(np.abs(B64[d] - b0[d]).max() for d in range(3))

    def err(s):
        return max(np.abs(s.download(5 + d).astype(np.float64) - B64[d]).max() for d in range(3)) / pert

    dp, dn, dc = divn(sp), divn(sn), divn(sc)
    ep, en, ec = err(sp), err(sn), err(sc)
    assert div0 < 3e-7
    assert dp > 10 * div0 and dn > 10 * div0
    assert dc < 3 * div0
    assert en > 3 * ec and ep > 3 * ec
    assert ec < 0.05
    assert np.abs(sc.ct_slot(18)).max() <= 2.1e-6      # remainders within half an ulp of the faces
    assert not np.any(sn.ct_slot(18)) and not np.any(sp.ct_slot(19))
    bxf = sc.ct_slot(0).reshape(N, N, N)
    assert np.array_equal((np.float32(0.5) * (bxf + np.roll(bxf, -1, axis=2))).reshape(-1), sc.download(5))
