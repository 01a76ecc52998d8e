"""Pins the CPU oracle before anything is checked against it.

* two independent restatements (per-cell C++ and whole-grid numpy) agree to round-off;
* closed forms and invariants (flux consistency, rotation equivariance, conservation);
* the reference's own recorded known answers (docs/data/validation-cpu.log) for the
  Alfven convergence study, low-beta Alfven, Brio-Wu, Orszag-Tang and the field loop;
* CT: the reference's fork-port cross-check (test/inherited/test_ct.jl:209-241) and its
  divergence / faces-from-cells invariants (test_ct.jl:243-334).
"""
import json
import os

import numpy as np
import pytest

import face_np as F
import oracle as O
from pmfmhd import problems as P

G = 5.0 / 3.0


def rand_prim(rng, n, rho_lo=0.2):
    q = np.empty((n, 9))
    q[:, 0] = rng.uniform(rho_lo, 2.0, n)
    q[:, 1:4] = rng.normal(0, 0.7, (n, 3))
    q[:, 4] = rng.uniform(0.05, 2.0, n)
    q[:, 5:8] = rng.normal(0, 0.8, (n, 3))
    q[:, 8] = rng.normal(0, 0.1, n)
    return q


@pytest.mark.parametrize("solver", ["hlld", "hll", "llf"])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_riemann_two_restatements_agree(solver, d):
    rng = np.random.default_rng(10 + d)
    n = 4000
    L, R = rand_prim(rng, n), rand_prim(rng, n)
    # floors and the LLF switch (glm_physics.jl:320-324) must be exercised
    L[:50, 0] = 1e-8
    R[50:100, 4] = -1.0
    L[100:150, 0] = 5e-5
    ch = 1.7
    Fc = O.riemann_batch(L, R, d, ch, riemann=solver)
    with np.errstate(all="ignore"):
        Fn = F.riemann(L.T, R.T, d, ch, G, solver=solver).T
    ok = np.isfinite(Fc).all(axis=1) & np.isfinite(Fn).all(axis=1)
    assert ok.mean() > 0.95
    scale = np.maximum(np.abs(Fc[ok]).max(axis=0), 1.0)
    assert np.max(np.abs(Fc[ok] - Fn[ok]) / scale) < 1e-12


@pytest.mark.parametrize("recon", ["plm", "ppm", "plm_minmod"])
def test_hancock_two_restatements_agree(recon):
    rng = np.random.default_rng(3)
    n = 2000
    M = np.stack([rand_prim(rng, n, rho_lo=0.5) for _ in range(7)], axis=1)
    M[:, 1:] = M[:, :1] + 0.1 * (M[:, 1:] - M[:, :1])  # smooth neighbourhoods
    dtdx = 0.15
    mh, dL, dR = O.predict_batch(M, dtdx, recon=recon)
    q = M.transpose(1, 2, 0)  # (7, 9, n)
    dLn, dRn = [], []
    for d in range(3):
        a, b = F.recon_offsets(q[1 + 2 * d], q[0], q[2 + 2 * d], recon)
        dLn.append(a)
        dRn.append(b)
    mhn = F.cons2prim(F.hancock(q[0], dLn, dRn, dtdx, G), G, 1e-6, 1e-7)
    assert np.max(np.abs(mh - mhn.T)) < 1e-13
    assert np.max(np.abs(dL - np.stack(dLn, 0).transpose(2, 0, 1))) < 1e-15
    assert np.max(np.abs(dR - np.stack(dRn, 0).transpose(2, 0, 1))) < 1e-15


def smooth_state(n, seed=0):
    rng = np.random.default_rng(seed)
    dims = (n, n, n)
    i, j, k = P._grid(dims)
    x, y, z = [2 * np.pi * (a + 0.5) / n for a in (i, j, k)]
    ph = rng.uniform(0, 2 * np.pi, 9)
    q = np.stack([1 + 0.2 * np.sin(x + ph[0]) * np.cos(y), 0.3 * np.sin(z + ph[1]), 0.3 * np.sin(x + ph[2]),
                  0.3 * np.sin(y + ph[3]), 1 + 0.3 * np.cos(z + ph[4]) * np.sin(x), 0.5 + 0.2 * np.sin(y + ph[5]),
                  0.3 * np.sin(z + ph[6]), 0.3 * np.cos(x + ph[7]), 0.05 * np.sin(x + y + ph[8])])
    return P.prim2cons(q, G)


@pytest.mark.parametrize("recon", ["plm", "ppm", "plm_minmod"])
@pytest.mark.parametrize("solver", ["hlld", "hll", "llf"])
def test_full_step_two_restatements_agree(recon, solver):
    n = 8
    U0 = smooth_state(n, seed=1)
    s = O.OracleState((n, n, n), 1.0 / n, riemann=solver, recon=recon)
    s.set_conserved(U0)
    dt, smax = s.compute_dt(0.3)
    ch = smax
    decay = np.exp(-O.GLM_CR * ch * dt / s.dx)
    s.step_ref(dt, ch, decay)
    Un = F.step_periodic(U0, (n, n, n), dt / s.dx, ch, decay, recon=recon, solver=solver)
    err = np.abs(s.U - Un).max(axis=1) / np.maximum(np.abs(Un).max(axis=1), 1e-3)
    assert err.max() < 1e-13
    assert np.abs(s.U - U0).max() > 1e-4  # the state moved


def test_flux_consistency_hlld_of_equal_states():
    """HLLD(U,U) is the physical flux; the GLM slots carry psi* = psi, ch^2 bn (closed form)."""
    rng = np.random.default_rng(5)
    Q = rand_prim(rng, 500, rho_lo=0.5)
    ch = 2.3
    for d in range(3):
        Fx = O.riemann_batch(Q, Q, d, ch)
        qr = Q[:, F.ROT_TO[d]]
        phys = F.phys_flux_x(qr.T, G).T
        phys[:, 5] = qr[:, 8]
        phys[:, 8] = ch * ch * qr[:, 5]
        phys = phys[:, F.ROT_FROM[d]]
        assert np.max(np.abs(Fx - phys)) < 1e-12


def test_rotation_equivariance_of_step():
    """A cyclically permuted grid advances to the permuted result (glm_physics.jl:36-45)."""
    n = 8
    U = smooth_state(n, seed=2).reshape(9, n, n, n)  # (var, z, y, x)
    # relabel axes x->y->z->x: new(x', y', z') field
    Up = np.empty_like(U)
    perm = [0, 3, 1, 2, 4, 7, 5, 6, 8]  # new vx = old vz, new vy = old vx, new vz = old vy
    for v in range(9):
        Up[v] = np.transpose(U[perm[v]], (1, 2, 0))
    a = O.OracleState((n, n, n), 1.0 / n)
    b = O.OracleState((n, n, n), 1.0 / n)
    a.set_conserved(U.reshape(9, -1))
    b.set_conserved(Up.reshape(9, -1))
    dt, smax = a.compute_dt(0.3)
    for s in (a, b):
        s.step(dt, ch=smax)
    Ua = a.U.reshape(9, n, n, n)
    Ub = b.U.reshape(9, n, n, n)
    for v in range(9):
        assert np.max(np.abs(Ub[v] - np.transpose(Ua[perm[v]], (1, 2, 0)))) < 1e-14


def test_conservation_periodic():
    n = 8
    s = O.OracleState((n, n, n), 1.0 / n)
    s.set_conserved(smooth_state(n, seed=4))
    t0 = s.conserved_totals()
    s.evolve(0.05, cfl=0.3)
    t1 = s.conserved_totals()
    for k in ("mass", "momx", "momy", "momz", "energy", "Bx", "By", "Bz"):
        assert abs(t1[k] - t0[k]) <= 1e-13 * max(1.0, abs(t0[k]))


# --- the reference's recorded known answers (docs/data/validation-cpu.log) ---------------
def alfven_L1(N, recon="plm", amp=0.1, p0=0.1, tfinal=0.5, cfl=0.4, glm_cr=O.GLM_CR, dtype=np.float32):
    """test/inherited/runtests.jl:6-25."""
    dims = (N, 4, 4)
    s = O.OracleState(dims, 1.0 / N, dtype=dtype, recon=recon)
    s.set_conserved(P.alfven_wave(dims, 1.0 / N, amp=amp, B0=1.0, p0=p0))
    tot0 = s.conserved_totals()
    t, n = s.evolve(tfinal, cfl=cfl, glm_cr=glm_cr)
    tot1 = s.conserved_totals()
    x = (np.arange(N) + 0.5) / N
    L1 = np.mean(np.abs(s.U[6][:N].astype(np.float64) - P.alfven_By_exact(x, t, 1.0, amp=amp)))
    return L1, n, abs(tot1["mass"] - tot0["mass"]) / tot0["mass"], abs(tot1["energy"] - tot0["energy"]) / tot0["energy"]


# validation-cpu.log:8-19 (Float32 @fastmath KA CPU reference, 3 significant figures),
# transcribed in tests/golden/known_answers.json
KNOWN = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "known_answers.json")))
RECORDED = {r: tuple((e["L1"], e["steps"]) for e in KNOWN["alfven_circular"][r]) for r in ("plm", "ppm")}


@pytest.mark.parametrize("recon", ["plm", "ppm"])
def test_alfven_convergence_matches_recorded(recon):
    L1s = []
    for (ref_L1, ref_n), N in zip(RECORDED[recon], (32, 64, 128)):
        L1, n, dmass, denergy = alfven_L1(N, recon=recon)
        assert n == ref_n
        # float32 round-off (IEEE here, @fastmath in the reference) moves L1 by ~0.1% at
        # N<=64; at N=128 the error is small enough that round-off spreads it by ~2%
        # (float64 3.610e-5, IEEE float32 3.553e-5, recorded 3.586e-5)
        tol = 5e-3 if N < 128 else 2.5e-2
        assert abs(L1 - ref_L1) / ref_L1 < tol, (N, L1, ref_L1)
        assert dmass < 1e-4 and denergy < 1e-4
        L1s.append(L1)
    orders = [np.log2(L1s[k - 1] / L1s[k]) for k in (1, 2)]
    assert min(orders) > 1.7


def test_alfven_float64_matches_recorded_tighter():
    # the float64 restatement sits on the recorded 3-figure values where float32 round-off
    # is below the last recorded digit (N <= 64), and within 1% at N=128
    for recon in ("plm", "ppm"):
        for (ref_L1, _), N in zip(RECORDED[recon], (32, 64, 128)):
            L1, _, _, _ = alfven_L1(N, recon=recon, dtype=np.float64)
            assert abs(L1 - ref_L1) / ref_L1 < (1.5e-3 if N < 128 else 1e-2), (recon, N, L1)


def test_low_beta_alfven_matches_recorded():
    """runtests.jl:27-35; validation-cpu.log:5: L1 1.512e-4, 89 steps, dE 8.6e-8."""
    K = KNOWN["alfven_low_beta"]
    L1, n, _, denergy = alfven_L1(K["N"], amp=0.1, p0=1e-3, glm_cr=1.0)
    assert n == K["steps"]
    assert abs(L1 - K["L1"]) / K["L1"] < 1e-2
    assert denergy < 2e-4


def brio_wu_density(N, tfinal=0.1):
    dims = (N, 4, 4)
    s = O.OracleState(dims, 1.0 / N, dtype=np.float32, gamma=2.0, bcs=("outflow", "periodic", "periodic"))
    s.set_conserved(P.brio_wu(dims, gamma=2.0))
    t0 = s.conserved_totals()
    t, n = s.evolve(tfinal, cfl=0.4)
    t1 = s.conserved_totals()
    return s.U[0][:N].astype(np.float64), abs(t1["mass"] - t0["mass"]) / t0["mass"], n


def test_brio_wu_matches_recorded():
    """runtests.jl:65-91; validation-cpu.log:22: min 0.117, L1(256 vs 512) 2.057e-3."""
    r2, _, _ = brio_wu_density(256)
    r4, dm4, _ = brio_wu_density(512)
    r4d = 0.5 * (r4[0::2] + r4[1::2])
    L1 = np.mean(np.abs(r2 - r4d))
    K = KNOWN["brio_wu"]
    assert abs(L1 - K["L1_256_vs_512"]) / K["L1_256_vs_512"] < 2e-2
    assert abs(r2.min() - K["rho_min"]) < 1.5e-3
    assert r2[0] > 0.9 and r2[-1] < 0.2 and r2.max() < 1.02
    assert dm4 < 1e-4


@pytest.mark.slow
def test_orszag_tang_matches_recorded():
    """runtests.jl:94-122; validation-cpu.log:25: 333 steps, max|divB|dx/Brms = 2.79."""
    N = 128
    dims = (N, N, 1)
    s = O.OracleState(dims, 1.0 / N, dtype=np.float32)
    s.set_conserved(P.orszag_tang(dims, 1.0 / N))
    t0 = s.conserved_totals()
    t, n = s.evolve(0.5, cfl=0.4)
    t1 = s.conserved_totals()
    h = s.U.astype(np.float64)
    brms = np.sqrt(np.mean(h[5] ** 2 + h[6] ** 2 + h[7] ** 2))
    norm = s.max_divb() * float(s.dx) / brms
    assert n == KNOWN["orszag_tang"]["steps"]
    assert abs(norm - KNOWN["orszag_tang"]["max_divb_dx_over_brms"]) < 0.05
    assert abs(t1["mass"] - t0["mass"]) / t0["mass"] < 1e-4
    assert h[0].min() > 0


@pytest.mark.slow
def test_field_loop_matches_recorded():
    """runtests.jl:486-509; validation-cpu.log:86: 734 steps, L1 0.1874, retention 0.895."""
    N = 64
    dims = (N, N, 4)
    s = O.OracleState(dims, 1.0 / N, dtype=np.float32)
    s.set_conserved(P.field_loop(dims, 1.0 / N))
    h0 = s.U.astype(np.float64)
    bnorm = np.abs(h0[5]).sum() + np.abs(h0[6]).sum()
    emag0 = (h0[5] ** 2).sum() + (h0[6] ** 2).sum()
    assert s.max_divb() < 2e-6
    t, n = s.evolve(1.0, cfl=0.4, glm_ch_fac=2.0, glm_cr=1.0)
    h1 = s.U.astype(np.float64)
    l1 = (np.abs(h1[5] - h0[5]).sum() + np.abs(h1[6] - h0[6]).sum()) / bnorm
    emag1 = (h1[5] ** 2).sum() + (h1[6] ** 2).sum()
    divn = s.max_divb() * float(s.dx) / np.sqrt(emag1 / h1[5].size)
    K = KNOWN["field_loop"]
    assert n == K["steps"]
    assert abs(l1 - K["L1"]) < 2e-3
    assert abs(emag1 / emag0 - K["emag_retention"]) < 2e-3
    assert divn < 0.5


# --- constrained transport ----------------------------------------------------------------
def ct_test_potential(N, amp=0.15):
    """test_ct.jl:147-163."""
    dx = 1.0 / N
    tp = 2 * np.pi
    k, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    xc, xe = (i + 0.5) * dx, i * dx
    yc, ye = (j + 0.5) * dx, j * dx
    zc, ze = (k + 0.5) * dx, k * dx
    Ax = amp * (np.sin(tp * ye + 0.3) * np.cos(tp * ze) + 0.5 * np.sin(tp * (ye + ze) + 1.0) / tp) / tp
    Ay = amp * (np.sin(tp * ze + 1.1) * np.cos(tp * xe + 0.4) + 0.4 * np.cos(tp * (xe - ze)) / tp) / tp
    Az = amp * (np.sin(tp * xe + 0.7) * np.cos(tp * ye + 0.2) + 0.3 * np.sin(tp * (xe + ye) + 0.5) / tp) / tp
    return Ax, Ay, Az


def ct_test_hydro(N, gamma):
    """test_ct.jl:165-178 (magnetic energy is added by the face install)."""
    tp = 2 * np.pi
    k, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    x, y, z = (i + 0.5) / N, (j + 0.5) / N, (k + 0.5) / N
    rho = 1 + 0.2 * np.sin(tp * x + 0.3) * np.cos(tp * y) + 0.1 * np.sin(tp * (y + z) + 0.9)
    vx = 0.3 * np.sin(tp * y + 0.1) + 0.1 * np.cos(tp * (x + z))
    vy = 0.3 * np.sin(tp * z + 0.8) + 0.1 * np.sin(tp * (x - y) + 0.2)
    vz = 0.3 * np.sin(tp * x + 0.5) + 0.1 * np.cos(tp * (y - z) + 0.6)
    p = 1 + 0.15 * np.cos(tp * (x + y) + 0.4) + 0.1 * np.sin(tp * z)
    U = np.zeros((9, N ** 3))
    U[0] = rho.reshape(-1)
    U[1] = (rho * vx).reshape(-1)
    U[2] = (rho * vy).reshape(-1)
    U[3] = (rho * vz).reshape(-1)
    U[4] = (p / (gamma - 1) + 0.5 * rho * (vx * vx + vy * vy + vz * vz)).reshape(-1)
    return U


def test_ct_matches_fork_port():
    """test_ct.jl:209-241: HLL, PLM MonCen, float64, N=12, 5 steps, rel max diff < 1e-10."""
    N = 12
    s = O.OracleState((N, N, N), 1.0 / N, riemann="hll", recon="plm", mhd_solver="ct", smallr=F.FK_SMALLR,
                      pfl=F.FK_SMALLP)
    s.set_conserved(ct_test_hydro(N, G))
    s.ct_faces_from_potential(*ct_test_potential(N))
    assert s.max_divb_face() * float(s.dx) < 1e-13
    U = s.U[:5].reshape(5, N, N, N).copy()
    bf = [s.ct[d].reshape(N, N, N).copy() for d in range(3)]
    bxf0 = bf[0].copy()
    dt, _ = s.compute_dt(0.3)
    for _ in range(5):
        s.step(dt)
        F.fork_ct_step(U, bf[0], bf[1], bf[2], dt, 1.0 / N, G)

    def rel(a, b):
        return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)

    errs = [rel(s.U[t].reshape(N, N, N), U[t]) for t in range(5)]
    ef = [rel(s.ct[d].reshape(N, N, N), bf[d]) for d in range(3)]
    assert max(errs) < 1e-10 and max(ef) < 1e-10, (errs, ef)
    assert rel(s.ct[0].reshape(N, N, N), bxf0) > 1e-3
    assert s.max_divb_face() * float(s.dx) < 1e-12


@pytest.mark.parametrize("dims", [(16, 16, 16), (12, 16, 8), (15, 9, 10)])
def test_ct_faces_from_cells_divergence_free(dims):
    """test_ct.jl:243-281 (white noise plus Nyquist content)."""
    nx, ny, nz = dims
    s = O.OracleState(dims, 1.0 / max(dims), mhd_solver="ct")
    rng = np.random.default_rng(7)
    i, j, k = P._grid(dims)
    U = np.zeros((9, s.nc))
    U[0] = 1
    U[4] = 1.5
    U[5] = 0.3 + 0.1 * np.sin(2 * np.pi * (j + 1) / ny) + 0.05 * (rng.random(s.nc) - 0.5) + 0.02 * (-1.0) ** (i + j + k)
    U[6] = 0.1 * np.cos(2 * np.pi * (k + 1) / nz + 0.3) + 0.05 * (rng.random(s.nc) - 0.5) + 0.02 * (-1.0) ** (i + j + k)
    U[7] = 0.1 * np.sin(2 * np.pi * (i + 1) / nx + 0.1) + 0.05 * (rng.random(s.nc) - 0.5) + 0.02 * (-1.0) ** (i + j + k)
    s.set_conserved(U)
    info = s.ct_faces_from_cells()
    assert info["max_divb_face_dx_over_brms"] < 1e-13
    assert info["cell_b_rel_l2_change"] < 0.3
    assert np.all(s.U[8] == 0)


def test_ct_divergence_preserved_under_evolution():
    """test_ct.jl:305-334 (float64, N=24 reduced to N=12 for time)."""
    N = 12
    s = O.OracleState((N, N, N), 1.0 / N, riemann="hlld", mhd_solver="ct")
    tp = 2 * np.pi
    i, j, k = P._grid((N, N, N))
    x, y, z = (i + 0.5) / N, (j + 0.5) / N, (k + 0.5) / N
    rho = 1 + 0.2 * np.sin(tp * x) * np.cos(tp * z)
    v = [0.5 * np.sin(tp * y), 0.5 * np.sin(tp * z + 0.4), 0.5 * np.sin(tp * x + 0.9)]
    b = [0.4 * np.sin(tp * y + 0.2) + 0.2 * np.cos(tp * z), 0.4 * np.sin(tp * z) + 0.2 * np.cos(tp * x + 0.5),
         0.4 * np.sin(tp * x + 1.0) + 0.2 * np.cos(tp * y)]
    s.set_conserved(P.prim2cons(np.stack([rho, *v, 0.6 * np.ones_like(x), *b, 0 * x]), G))
    info = s.ct_faces_from_cells()
    divs = []
    s.evolve(0.1, cfl=0.3, callback=lambda st, t, n: divs.append(st.max_divb_face() * float(st.dx)))
    assert info["max_divb_face_dx_over_brms"] < 1e-12
    assert max(divs) < 1e-12
    assert np.all(s.U[8] == 0)


def test_reflecting_wall_equals_mirrored_periodic_domain():
    """boundary.jl:17-31: a reflecting wall is a mirror. A half domain with reflecting x walls
    advances exactly like the left half of a periodic domain holding the state and its mirror
    image (normal velocity and normal field flipped)."""
    nx, ny, nz = 8, 6, 5
    U = smooth_state(8, seed=9).reshape(9, 8, 8, 8)[:, :nz, :ny, :nx]
    U = np.ascontiguousarray(U)
    flip = np.array([1, -1, 1, 1, 1, -1, 1, 1, 1.0])[:, None, None, None]
    Uf = np.concatenate([U, (U * flip)[:, :, :, ::-1]], axis=3)  # periodic in x with period 2 nx
    a = O.OracleState((nx, ny, nz), 1.0 / 8, bcs=("reflecting", "periodic", "periodic"), recon="ppm")
    b = O.OracleState((2 * nx, ny, nz), 1.0 / 8, recon="ppm")
    a.set_conserved(U.reshape(9, -1))
    b.set_conserved(Uf.reshape(9, -1))
    dt, smax = b.compute_dt(0.3)
    for _ in range(3):
        a.step(dt, ch=smax)
        b.step(dt, ch=smax)
    Ub = b.U.reshape(9, nz, ny, 2 * nx)[:, :, :, :nx].reshape(9, -1)
    assert np.max(np.abs(a.U - Ub)) < 1e-13


def test_llf_consistency_and_minmod_convergence():
    """LLF(U,U) is the physical flux (glm_physics.jl:239-244); minmod PLM converges at
    second order on the reference's Alfven study (runtests.jl:6-25)."""
    rng = np.random.default_rng(6)
    Q = rand_prim(rng, 300, rho_lo=0.5)
    for d in range(3):
        Fx = O.riemann_batch(Q, Q, d, 1.3, riemann="llf")
        qr = Q[:, F.ROT_TO[d]]
        phys = F.phys_flux_x(qr.T, G).T
        phys[:, 5] = qr[:, 8]
        phys[:, 8] = 1.3 * 1.3 * qr[:, 5]
        assert np.max(np.abs(Fx - phys[:, F.ROT_FROM[d]])) < 1e-12
    L1 = [alfven_L1(N, recon="plm_minmod", dtype=np.float64)[0] for N in (32, 64, 128)]
    orders = [np.log2(L1[k - 1] / L1[k]) for k in (1, 2)]
    assert min(orders) > 1.6, orders
    # minmod is more diffusive than MonCen on the same wave (validation-cpu.log:8-12)
    assert L1[0] > 6.919e-4


# ---------------------------------------------------------------------------
# Closed-form pins of the dissipation speeds, the floors and the HLLD limits. Each
# expected value is computed here from textbook formulas, not from the oracle's code.

def _cf_textbook(rho, p, B, bn, g=G):
    """Fast magnetosonic speed: c_f^2 = (a^2 + b^2 + sqrt((a^2 + b^2)^2 - 4 a^2 b_n^2)) / 2
    with a^2 = g p / rho, b^2 = |B|^2 / rho, b_n^2 = bn^2 / rho."""
    a2, b2, bn2 = g * p / rho, np.dot(B, B) / rho, bn * bn / rho
    return np.sqrt(0.5 * (a2 + b2 + np.sqrt((a2 + b2) ** 2 - 4 * a2 * bn2)))


def _rest(rho, p, B, psi=0.0):
    return np.array([rho, 0.0, 0.0, 0.0, p, B[0], B[1], B[2], psi])


def _perm(q, d):
    """The state whose direction-d frame is q's x frame: normal components go to axis d."""
    out = q.copy()
    for base in (1, 5):
        comp = q[base:base + 3]
        out[base:base + 3] = np.roll(comp, d)
    return out


@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("ch,ch_dominates", [(10.0, True), (0.05, False)])
def test_llf_dissipation_of_a_resting_density_jump(d, ch, ch_dominates):
    """glm_physics.jl:239-244. Two resting states with equal p and B and a density jump:
    the central part of every flux is the same on both sides, so F = F(U) - smax/2 (U_R - U_L)
    with smax = max(c_f,L, c_f,R, ch). Only rho differs, so F_rho = -smax (rho_R - rho_L)/2
    (negative: mass moves from the dense side), F_E = 0, the normal momentum flux is the
    total pressure minus Bn^2, F_psi = ch^2 Bn and F_Bn = psi* = 0."""
    B = np.array([0.4, -0.3, 0.2])
    p, rl, rr = 0.7, 1.0, 2.5
    L, R = _perm(_rest(rl, p, B), d), _perm(_rest(rr, p, B), d)
    Fx = O.riemann_batch(L[None], R[None], d, ch, riemann="llf")[0]
    cf = max(_cf_textbook(rl, p, B, B[0]), _cf_textbook(rr, p, B, B[0]))
    smax = max(cf, ch)
    assert (smax == ch) == ch_dominates
    n = [1 + d, 5 + d]  # normal momentum and normal field slots in the grid frame
    t = [1 + (d + 1) % 3, 1 + (d + 2) % 3]
    want = np.zeros(9)
    want[0] = -0.5 * smax * (rr - rl)
    want[n[0]] = p + 0.5 * B @ B - B[0] ** 2
    want[t[0]], want[t[1]] = -B[0] * B[1], -B[0] * B[2]
    want[8] = ch * ch * B[0]
    np.testing.assert_allclose(Fx, want, rtol=1e-14, atol=1e-15)
    assert Fx[0] < 0


@pytest.mark.parametrize("d", [0, 1, 2])
def test_hll_dissipation_of_a_resting_density_jump(d):
    """glm_physics.jl:245-252: with S_L = -S_R = -max c_f, HLL gives F_rho = -c_f (rho_R - rho_L)/2
    and does not take ch into its speeds."""
    B = np.array([0.4, -0.3, 0.2])
    p, rl, rr = 0.7, 1.0, 2.5
    L, R = _perm(_rest(rl, p, B), d), _perm(_rest(rr, p, B), d)
    F = O.riemann_batch(L[None], R[None], d, 30.0, riemann="hll")[0]
    cf = max(_cf_textbook(rl, p, B, B[0]), _cf_textbook(rr, p, B, B[0]))
    assert F[0] == pytest.approx(-0.5 * cf * (rr - rl), rel=1e-14)
    assert F[4] == pytest.approx(0.0, abs=1e-15)


def test_cons2prim_floors_density_before_velocity_and_kinetic_energy():
    """glm_physics.jl:11-18 through the predictor of a uniform neighbourhood (no slopes, so
    mh = cons2prim(prim2cons(q))): a density below smallr is raised to smallr first, so
    v = m / smallr and p = p + (g-1) rho v^2 (1 - rho/smallr) / 2; a state whose internal
    energy is negative gets p = pfl."""
    smallr, pfl = 1e-6, 1e-7
    rho, v, p = 2.5e-7, np.array([0.3, -0.2, 0.1]), 0.02
    B = np.array([0.1, 0.05, -0.2])
    q_low = np.concatenate([[rho], v, [p], B, [0.01]])
    q_neg = np.array([1.2, 0.4, 0.0, -0.3, -0.5, 0.2, 0.1, 0.3, -0.02])
    M = np.stack([np.repeat(q[None], 7, axis=0) for q in (q_low, q_neg)])
    mh, _, _ = O.predict_batch(M, 0.2, recon="plm", smallr=smallr, pfl=pfl)
    np.testing.assert_allclose(mh[0, 0], smallr, rtol=0)
    np.testing.assert_allclose(mh[0, 1:4], rho * v / smallr, rtol=1e-14)
    want_p = p + (G - 1) * 0.5 * rho * (v @ v) * (1 - rho / smallr)
    np.testing.assert_allclose(mh[0, 4], want_p, rtol=1e-12)
    np.testing.assert_allclose(mh[0, 5:9], q_low[5:9], rtol=1e-15)
    assert mh[1, 4] == pfl
    np.testing.assert_allclose(mh[1, [0, 1, 2, 3, 5, 6, 7, 8]], q_neg[[0, 1, 2, 3, 5, 6, 7, 8]], rtol=1e-15,
                               atol=1e-17)


@pytest.mark.parametrize("riemann", ["llf", "hll", "hlld"])
def test_riemann_floors_both_face_states(riemann):
    """glm_physics.jl:320-321: the face states are floored again before the solve. Two resting,
    field-free states whose densities and pressures are both below the floors are the same
    state after flooring, so every solver returns the physical flux of (smallr, 0, pfl):
    no mass flux and a normal momentum flux of exactly pfl (unfloored it would be -1.5)."""
    smallr, pfl = 1e-6, 1e-7
    L = np.array([1e-9, 0, 0, 0, -1.0, 0, 0, 0, 0])
    R = np.array([3e-7, 0, 0, 0, -2.0, 0, 0, 0, 0])
    for d in range(3):
        F = O.riemann_batch(L[None], R[None], d, 1.0, riemann=riemann, smallr=smallr, pfl=pfl, llf_dmin=0.0)[0]
        want = np.zeros(9)
        want[1 + d] = pfl
        np.testing.assert_allclose(F, want, rtol=1e-14, atol=1e-20)  # pfl * c_f ~ 3e-8


def _hllc_normal_field(L, R, g=G):
    """Textbook HLLC (Toro 2009, sec. 10.4; Batten et al. 1997 star pressure) for the x
    flux of two states whose field is purely normal (By = Bz = 0, equal Bn). Such a field
    only adds Bn^2/2 to the pressure and to the energy and is carried unchanged, so the
    MHD flux is the hydro HLLC flux of (rho, rho u, E - Bn^2/2) with the fast speed
    max(a, |Bn|/sqrt(rho)), the momentum flux lowered by Bn^2/2, F_E the hydro one,
    and no transverse fluxes."""
    bn = L[5]
    out = []
    for q in (L, R):
        rho, u, p = q[0], q[1], q[4]
        e_h = p / (g - 1) + 0.5 * rho * u * u
        cf = max(np.sqrt(g * p / rho), abs(bn) / np.sqrt(rho))
        out.append((rho, u, p, e_h, cf))
    (rl, ul, pl, el, cfl), (rr, ur, pr, er, cfr) = out
    SL = min(min(ul, ur) - max(cfl, cfr), 0.0)
    SR = max(max(ul, ur) + max(cfl, cfr), 0.0)
    SM = (pr - pl + rl * ul * (SL - ul) - rr * ur * (SR - ur)) / (rl * (SL - ul) - rr * (SR - ur))
    pstar = pl + rl * (SL - ul) * (SM - ul)

    def flux(rho, u, p, e):
        return np.array([rho * u, rho * u * u + p, (e + p) * u])

    def star(rho, u, p, e, S):
        f = (S - u) / (S - SM)
        return np.array([rho * f, rho * f * SM, ((S - u) * e - p * u + pstar * SM) / (S - SM)])

    if SM >= 0:
        F = flux(rl, ul, pl, el) + SL * (star(rl, ul, pl, el, SL) - np.array([rl, rl * ul, el]))
    else:
        F = flux(rr, ur, pr, er) + SR * (star(rr, ur, pr, er, SR) - np.array([rr, rr * ur, er]))
    out = np.zeros(9)
    out[0], out[1], out[4] = F[0], F[1] - 0.5 * bn * bn, F[2]
    return out


def test_hlld_reduces_to_hllc_for_a_purely_normal_field():
    """glm_physics.jl:255-307 in the limit By = Bz = 0: the rotational discontinuities
    vanish, both star states coincide with the double-star ones and HLLD is HLLC. Random
    jumps in rho, u and p, including low-beta states (g p << Bn^2, where c_f is the
    Alfven speed and the fast and Alfven waves are degenerate)."""
    rng = np.random.default_rng(41)
    n = 400
    L = np.zeros((n, 9))
    R = np.zeros((n, 9))
    for q in (L, R):
        q[:, 0] = rng.uniform(0.2, 3.0, n)
        q[:, 1] = rng.normal(0, 1.0, n)
        q[:, 4] = rng.uniform(0.05, 2.0, n)
    bn = rng.normal(0, 1.0, n)
    bn[: n // 4] *= 20.0  # low beta
    L[:, 5] = R[:, 5] = bn
    F = O.riemann_batch(L, R, 0, 1.0, riemann="hlld", llf_dmin=0.0)
    for f in range(n):
        want = _hllc_normal_field(L[f], R[f])
        want[8] = 1.0 * bn[f]  # F_psi = ch^2 Bn* (psi = 0, equal Bn)
        scale = max(np.abs(want).max(), 1.0)
        assert np.max(np.abs(F[f] - want)) < 1e-12 * scale, f


def test_hlld_degenerate_star_is_the_physical_flux():
    """glm_physics.jl:255-263: a resting state with a purely normal field and g p << Bn^2
    makes the fast speed the Alfven speed, so denK = rho (S-u)(S-SM) - Bn^2 vanishes
    analytically on both sides. The solve must take the degenerate branch (no 0/0) and, the
    two sides being equal, return the physical flux: only the momentum flux p - Bn^2/2 and
    F_psi = ch^2 Bn. Also at the density floor (rho = smallr, Bn scaled to keep the
    Alfven speed), where denK is a difference of numbers of order 1e-6."""
    for rho, bn, p in ((1.0, 1.0, 1e-4), (1e-6, 1e-3, 1e-12), (2.0, -0.7, 3e-5)):
        q = np.array([rho, 0, 0, 0, p, bn, 0, 0, 0])
        for d in range(3):
            qd = _perm(q, d)
            F = O.riemann_batch(qd[None], qd[None], d, 2.0, riemann="hlld", llf_dmin=0.0, smallr=1e-6,
                                pfl=1e-13)[0]
            assert np.all(np.isfinite(F))
            want = np.zeros(9)
            want[1 + d] = p - 0.5 * bn * bn
            want[8] = 4.0 * bn
            np.testing.assert_allclose(F, want, rtol=1e-12, atol=1e-12 * max(abs(p), bn * bn))


def test_llf_switch_takes_low_density_and_low_pressure_faces():
    """glm_physics.jl:324-327: with min rho < llf_dmin (or llf_pmin > 0 and min p < llf_pmin)
    the HLLD dispatch returns the LLF flux. A resting contact (density jump, equal p, normal
    field) is exact for HLLD (no mass flux) and diffused by LLF (-smax/2 drho)."""
    B = np.array([0.3, 0.0, 0.0])
    L, R = _rest(5e-5, 1.0, B), _rest(2e-4, 1.0, B)
    ch = 0.5
    on = O.riemann_batch(L[None], R[None], 0, ch, riemann="hlld", llf_dmin=1e-4)[0]
    off = O.riemann_batch(L[None], R[None], 0, ch, riemann="hlld", llf_dmin=0.0)[0]
    cf = max(_cf_textbook(5e-5, 1.0, B, B[0]), _cf_textbook(2e-4, 1.0, B, B[0]), ch)
    assert on[0] == pytest.approx(-0.5 * cf * (2e-4 - 5e-5), rel=1e-13)
    assert off[0] == pytest.approx(0.0, abs=1e-15)
    L2, R2 = _rest(1.0, 1e-3, B), _rest(3.0, 1e-3, B)
    Lp = O.riemann_batch(L2[None], R2[None], 0, ch, riemann="hlld", llf_dmin=0.0, llf_pmin=1e-2)[0]
    cf = max(_cf_textbook(1.0, 1e-3, B, B[0]), _cf_textbook(3.0, 1e-3, B, B[0]), ch)
    assert Lp[0] == pytest.approx(-0.5 * cf * 2.0, rel=1e-13)
    assert O.riemann_batch(L2[None], R2[None], 0, ch, riemann="hlld", llf_dmin=0.0)[0][0] == pytest.approx(0, abs=1e-14)
