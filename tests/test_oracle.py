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
