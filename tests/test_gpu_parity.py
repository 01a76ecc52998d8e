"""GPU parity: the HIP library (through its C ABI) against the CPU oracle on the same
seeded inputs.

Bars: float64 GPU vs float64 oracle to round-off (the kernel logic is the oracle's);
float32 GPU vs the float32 oracle <= 1e-5 relative L2 after one step; float32 GPU vs
the float64 oracle <= 1e-5 relative L2 after the 50-step 32^3 benchmark run (the
measured float32 round-off floor there is ~2e-6).
"""
import numpy as np
import pytest

import face_np as F
import oracle as O
import pmfmhd as M
from pmfmhd import problems as P

pytestmark = pytest.mark.gpu
G = 5.0 / 3.0


def relL2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def smooth_state(dims, seed=0):
    rng = np.random.default_rng(seed)
    i, j, k = P._grid(dims)
    x, y, z = [2 * np.pi * (a + 0.5) / n for a, n in zip((i, j, k), dims)]
    ph = rng.uniform(0, 2 * np.pi, 9)
    q = np.stack([1 + 0.2 * np.sin(x + ph[0]) * np.cos(y), 0.3 * np.sin(z + ph[1]), 0.3 * np.sin(x + ph[2]),
                  0.3 * np.sin(y + ph[3]), 1 + 0.3 * np.cos(z + ph[4]) * np.sin(x), 0.5 + 0.2 * np.sin(y + ph[5]),
                  0.3 * np.sin(z + ph[6]), 0.3 * np.cos(x + ph[7]), 0.05 * np.sin(x + y + ph[8])])
    return P.prim2cons(q, G)


def pair(dims, U, dtype=np.float64, **kw):
    g = M.allocate_state(dims, 1.0 / dims[0], dtype=dtype, **kw)
    g.set_conserved(U)
    okw = {k: v for k, v in kw.items() if k != "bc"}
    if "bc" in kw:
        okw["bcs"] = (kw["bc"],) * 3
    o = O.OracleState(dims, 1.0 / dims[0], dtype=dtype, **okw)
    o.set_conserved(U)
    return g, o


def rand_prim(rng, n):
    q = np.empty((n, 9))
    q[:, 0] = rng.uniform(0.2, 2.0, n)
    q[:, 1:4] = rng.normal(0, 0.7, (n, 3))
    q[:, 4] = rng.uniform(0.05, 2.0, n)
    q[:, 5:8] = rng.normal(0, 0.8, (n, 3))
    q[:, 8] = rng.normal(0, 0.1, n)
    return q


@pytest.mark.parametrize("solver", ["hlld", "hll", "llf"])
def test_riemann_flux_batch_matches_oracle(solver):
    rng = np.random.default_rng(1)
    n = 5000
    L, R = rand_prim(rng, n), rand_prim(rng, n)
    L[:40, 0] = 1e-8   # density floor
    R[40:80, 4] = -1   # pressure floor
    L[80:120, 0] = 5e-5  # LLF switch (llf_dmin)
    for dtype, tol in ((np.float64, 1e-12), (np.float32, 2e-5)):
        s = M.allocate_state((4, 4, 4), 0.25, dtype=dtype, riemann=solver)
        for d in range(3):
            Fg = M.riemann_flux(s, L, R, d, ch=1.7).astype(np.float64)
            Fo = O.riemann_batch(L.astype(dtype), R.astype(dtype), d, 1.7, dtype=dtype, riemann=solver)
            ok = np.isfinite(Fo).all(axis=1)
            scale = np.maximum(np.abs(Fo[ok]).max(axis=0), 1.0)
            assert np.max(np.abs(Fg[ok] - Fo[ok]) / scale) < tol, (dtype, d)


def corner_faces(seed=5):
    """Face pairs at the corners the CPU closed-form pins cover: resting density jumps
    (LLF/HLL dissipation), states below both floors, low-density and low-pressure faces
    (the LLF switch), purely normal fields (the HLLC limit of HLLD), exactly degenerate
    HLLD stars (g p << Bn^2 with equal sides, also at the density floor) and their
    neighbourhood (tiny transverse fields and jumps, so |denK| sits near 1e-12). Returns
    L, R and the index ranges of each group."""
    rng = np.random.default_rng(seed)
    groups = {}
    Ls, Rs = [], []

    def add(name, L, R):
        start = sum(len(x) for x in Ls)
        Ls.append(np.asarray(L, float))
        Rs.append(np.asarray(R, float))
        groups[name] = slice(start, start + len(L))

    n = 64
    rest = np.zeros((n, 9))
    rest[:, 0] = rng.uniform(0.2, 3.0, n)
    rest[:, 4] = rng.uniform(0.05, 2.0, n)
    rest[:, 5:8] = rng.normal(0, 0.6, (n, 3))
    restR = rest.copy()
    restR[:, 0] = rng.uniform(0.2, 3.0, n)
    add("rest_jump", rest, restR)
    fl = np.zeros((n, 9))
    fl[:, 0] = 10.0 ** rng.uniform(-10, -6.2, n)
    fl[:, 4] = -rng.uniform(0.1, 2, n)
    flR = fl.copy()
    flR[:, 0] = 10.0 ** rng.uniform(-10, -6.2, n)
    flR[:, 1:4] = rng.normal(0, 0.1, (n, 3))
    add("floors", fl, flR)
    sw = rand_prim(rng, n)
    sw[: n // 2, 0] = rng.uniform(1e-6, 1e-4, n // 2)
    sw[n // 2:, 4] = rng.uniform(1e-7, 1e-3, n - n // 2)
    add("switch", sw, rand_prim(rng, n))
    nf = np.zeros((n, 9))
    nf[:, 0] = rng.uniform(0.2, 3.0, n)
    nf[:, 1] = rng.normal(0, 1, n)
    nf[:, 4] = rng.uniform(0.05, 2.0, n)
    nfR = nf.copy()
    nfR[:, 0] = rng.uniform(0.2, 3.0, n)
    nfR[:, 1] = rng.normal(0, 1, n)
    nfR[:, 4] = rng.uniform(0.05, 2.0, n)
    nf[:, 5] = nfR[:, 5] = rng.normal(0, 3.0, n)
    add("normal_field", nf, nfR)
    dg = []
    for rho, bn, p in ((1.0, 1.0, 1e-4), (1e-6, 1e-3, 1e-12), (2.0, -0.7, 3e-5), (0.5, 2.0, 1e-3)):
        dg.append([rho, 0.3, 0, 0, p, bn, 0, 0, 0])
    add("degenerate", dg, dg)
    nd, ndR = [], []
    for rho, bn, p in ((1.0, 1.0, 1e-4), (1e-6, 1e-3, 1e-12), (2.0, -0.7, 3e-5)):
        va = abs(bn) / np.sqrt(rho)
        for bt in (1e-9, 1e-7, 1e-5, 1e-3):
            for du in (0.0, 1e-6, 1e-3):
                a = [rho, 0.1 * va, 0.05 * va, 0, p, bn, bt * abs(bn), -0.5 * bt * abs(bn), 0]
                b = list(a)
                b[1] += du * va
                b[6] *= 1.0 + du
                nd.append(a)
                ndR.append(b)
    add("near_degenerate", nd, ndR)
    return np.concatenate(Ls), np.concatenate(Rs), groups


@pytest.mark.parametrize("solver", ["hlld", "hll", "llf"])
def test_riemann_flux_batch_corners_match_oracle(solver):
    """pmf_riemann_flux_batch at the closed-form-pinned corners (floors, the LLF switch on
    density and on pressure, resting jumps, the normal-field HLLC limit, exactly and nearly
    degenerate HLLD stars) with llf_dmin 0 (so low-density faces reach HLLD) and with the
    defaults plus llf_pmin: float64 against the float64 oracle to 1e-12; float32 no
    further from the float64 oracle than max(2e-5, 4x) the float32 oracle's own distance
    from it (itself below 1e-4). Two groups are ill-conditioned in float32 in the
    reference's algorithm itself: near-degenerate stars, whose |denK| is float32 rounding
    noise, and HLLD on densities below 1e-4 once the LLF switch is off; there the two
    float32 evaluations round differently and 1e-3 of the flux scale is required."""
    L, R, groups = corner_faces()
    for kw in (dict(llf_dmin=0.0), dict(llf_dmin=1e-4, llf_pmin=1e-3)):
        for ch in (0.05, 2.0, 10.0):
            for d in range(3):
                F64 = O.riemann_batch(L, R, d, ch, riemann=solver, **kw)
                assert np.all(np.isfinite(F64))
                s = M.allocate_state((4, 4, 4), 0.25, dtype=np.float64, riemann=solver, **kw)
                G64 = M.riemann_flux(s, L, R, d, ch=ch)
                scale = np.maximum(np.abs(F64).max(axis=0), 1e-30)
                assert np.max(np.abs(G64 - F64) / scale) < 1e-12, (kw, ch, d)
                L32, R32 = L.astype(np.float32), R.astype(np.float32)
                F32 = O.riemann_batch(L32, R32, d, ch, dtype=np.float32, riemann=solver, **kw).astype(np.float64)
                s32 = M.allocate_state((4, 4, 4), 0.25, dtype=np.float32, riemann=solver, **kw)
                G32 = M.riemann_flux(s32, L32, R32, d, ch=ch).astype(np.float64)
                assert np.all(np.isfinite(G32))
                scale = np.maximum(np.abs(F64).max(axis=0), 1.0)
                e_o = np.max(np.abs(F32 - F64) / scale, axis=1)  # the algorithm's own float32 error
                e_g = np.max(np.abs(G32 - F64) / scale, axis=1)
                ill = np.zeros(len(L), dtype=bool)
                ill[groups["near_degenerate"]] = True
                if kw["llf_dmin"] == 0.0:
                    ill[groups["switch"]] = np.minimum(L[groups["switch"], 0], R[groups["switch"], 0]) < 1e-4
                assert np.all(e_o[~ill] < 1e-4), (kw, ch, d)
                assert np.all(e_g[ill] < 1e-3), (kw, ch, d)
                bad = np.flatnonzero(~ill & (e_g > np.maximum(2e-5, 4 * e_o)))
                assert bad.size == 0, (kw, ch, d, bad, e_g[bad], e_o[bad])


@pytest.mark.parametrize("recon", ["plm", "ppm", "plm_minmod"])
@pytest.mark.parametrize("solver", ["hlld", "hll", "llf"])
def test_ref_step_float64_matches_oracle(recon, solver):
    dims = (12, 10, 8)
    U = smooth_state(dims, seed=3)
    g, o = pair(dims, U, recon=recon, riemann=solver)
    dt, smax = o.compute_dt(0.3)
    dtg, smg = M.compute_dt(g, 0.3)
    assert abs(smg - smax) <= 1e-14 * smax and abs(dtg - dt) <= 1e-14 * dt
    for _ in range(2):
        o.step(dt, ch=smax)
        M.step(g, dt, ch=smax)
    Ug = g.conserved()
    for v in range(9):
        assert np.max(np.abs(Ug[v] - o.U[v])) <= 1e-12 * max(np.abs(o.U[v]).max(), 1e-3), v


@pytest.mark.parametrize("bcs", [("outflow", "periodic", "reflecting"), ("reflecting", "outflow", "periodic"),
                                 ("outflow", "outflow", "outflow")])
def test_ref_step_boundaries_match_oracle(bcs):
    dims = (10, 9, 8)
    U = smooth_state(dims, seed=5)
    g = M.allocate_state(dims, 0.1, dtype=np.float64, bcs=bcs, recon="ppm")
    g.set_conserved(U)
    o = O.OracleState(dims, 0.1, bcs=bcs, recon="ppm")
    o.set_conserved(U)
    dt, smax = o.compute_dt(0.3)
    for _ in range(2):
        o.step(dt, ch=smax)
        M.step(g, dt, ch=smax)
    Ug = g.conserved()
    for v in range(9):
        assert np.max(np.abs(Ug[v] - o.U[v])) <= 1e-12 * max(np.abs(o.U[v]).max(), 1e-3), v
    assert abs(M.max_divb(g) - o.max_divb()) <= 1e-12 * max(o.max_divb(), 1e-12)


def test_config1_one_step_float32_vs_float32_oracle():
    """Acceptance bar: float32 GPU vs the float32 CPU run <= 1e-5 after one step."""
    n = 32
    U = P.config_state("bstatic_32", seed=1)
    g, o = pair((n, n, n), U, dtype=np.float32, recon="plm_minmod")
    dt, smax = o.compute_dt(0.2)
    o.step(float(dt), ch=smax)
    M.step(g, float(dt), ch=smax)
    Ug = g.conserved()
    for v in range(9):
        assert relL2(Ug[v], o.U[v]) <= 1e-5, v


def test_config1_50_steps_float32_vs_float64_oracle():
    """The 32^3 benchmark run (50 CFL steps, cfl 0.2): each run computes its own dt."""
    n = 32
    U = P.config_state("bstatic_32", seed=1)
    g = M.allocate_state((n, n, n), 1.0 / n, dtype=np.float32, recon="plm_minmod")
    g.set_conserved(U)
    o = O.OracleState((n, n, n), 1.0 / n, dtype=np.float64, recon="plm_minmod")
    o.set_conserved(U)
    for _ in range(50):
        dt, smax = o.compute_dt(0.2)
        o.step(float(dt), ch=smax)
        dtg, smg = M.compute_dt(g, 0.2)
        M.step(g, float(dtg), ch=smg)
        assert abs(float(dtg) - float(dt)) <= 1e-5 * float(dt)
    Ug = g.conserved()
    errs = [relL2(Ug[v], o.U[v]) for v in range(9)]
    assert max(errs) <= 1e-5, errs
    tg, to = M.conserved_totals(g), o.conserved_totals()
    assert abs(tg["mass"] - to["mass"]) <= 1e-6 * to["mass"]


def test_evolve_alfven_float32_matches_recorded():
    """runtests.jl:6-25 on the GPU: N=32, 48 steps, L1(By) ~ 6.919e-4 (validation-cpu.log:8)."""
    N = 32
    s = M.allocate_state((N, 4, 4), 1.0 / N, dtype=np.float32)
    M.init_alfven_wave(s, amp=0.1, B0=1.0, p0=0.1)
    t, n = M.evolve(s, 0.5, cfl=0.4)
    By = s.download(6)[:N].astype(np.float64)
    x = (np.arange(N) + 0.5) / N
    L1 = np.mean(np.abs(By - P.alfven_By_exact(x, t, 1.0, amp=0.1)))
    assert n == 48
    assert abs(L1 - 6.919e-4) / 6.919e-4 < 5e-3


def test_diagnostics_match_oracle():
    dims = (16, 12, 10)
    U = smooth_state(dims, seed=7)
    g, o = pair(dims, U)
    assert abs(M.max_wavespeed(g) - o.max_wavespeed()) <= 1e-14 * o.max_wavespeed()
    assert abs(M.max_divb(g) - o.max_divb()) <= 1e-13 * o.max_divb()
    tg, to = M.conserved_totals(g), o.conserved_totals()
    for k in to:
        assert abs(tg[k] - to[k]) <= 1e-13 * max(1.0, abs(to[k])), k
    # float64 L2 norm of the centred divergence against numpy
    l2, linf = M.divb_norm(g)
    nx, ny, nz = dims
    b = [U[5 + d].reshape(nz, ny, nx) for d in range(3)]
    div = ((np.roll(b[0], -1, 2) - np.roll(b[0], 1, 2)) + (np.roll(b[1], -1, 1) - np.roll(b[1], 1, 1)) +
           (np.roll(b[2], -1, 0) - np.roll(b[2], 1, 0))) * (dims[0] / 2.0)
    assert abs(l2 - np.sqrt(np.mean(div ** 2))) <= 1e-12 * l2
    assert abs(linf - np.abs(div).max()) <= 1e-12 * linf


def test_float32_totals_are_float64_accurate():
    n = 32
    U = P.config_state("bstatic_32", seed=2)
    g = M.allocate_state((n, n, n), 1.0 / n, dtype=np.float32)
    g.set_conserved(U)
    tg = M.conserved_totals(g)
    U32 = U.astype(np.float32).astype(np.float64)
    for v, k in enumerate(("mass", "momx", "momy", "momz", "energy", "Bx", "By", "Bz")):
        ref = np.sum(U32[v]) / n ** 3
        assert abs(tg[k] - ref) <= 1e-14 * max(1.0, abs(ref)), k


# --- constrained transport ---------------------------------------------------------------
def test_ct_float64_matches_oracle_and_fork_port():
    from test_oracle import ct_test_hydro, ct_test_potential
    N = 12
    kw = dict(riemann="hll", recon="plm", mhd_solver="ct", smallr=F.FK_SMALLR, pfl=F.FK_SMALLP)
    g = M.allocate_state((N, N, N), 1.0 / N, dtype=np.float64, **kw)
    o = O.OracleState((N, N, N), 1.0 / N, **kw)
    U = ct_test_hydro(N, G)
    A = ct_test_potential(N)
    g.set_conserved(U)
    o.set_conserved(U)
    M.ct_faces_from_potential(g, *A)
    o.ct_faces_from_potential(*A)
    assert M.max_divb_face(g) / N < 1e-13
    dt, _ = o.compute_dt(0.3)
    for _ in range(5):
        o.step(dt)
        M.step(g, dt)
    Ug = g.conserved()
    for v in range(9):
        assert np.max(np.abs(Ug[v] - o.U[v])) <= 1e-12 * max(np.abs(o.U[v]).max(), 1e-6), v
    for d in range(3):
        assert np.max(np.abs(g.ct_slot(d) - o.ct[d])) <= 1e-12 * np.abs(o.ct[d]).max()
    assert M.max_divb_face(g) / N < 1e-12
    assert np.all(Ug[8] == 0)
    # the cell-field pass folded the finished state's CFL speed (timestep.jl:8-12)
    _, sm = M.compute_dt(g, 0.3)
    assert abs(sm - o.max_wavespeed()) <= 1e-14 * o.max_wavespeed()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ct_step_cfl_fold_equals_a_fresh_reduction(dtype):
    """After a CT step the library keeps the CFL speed its cell-field pass reduced; it is
    the speed a separate pass over the new state gives (to the last bit but for FMA
    contraction, which may differ between the two kernels), and a speed of other arrays
    does not leave a stale value."""
    tol = 4e-7 if dtype == np.float32 else 1e-15
    N = 24
    s = M.allocate_state((N, N, N), 1.0 / N, dtype=dtype, riemann="hlld", recon="ppm", mhd_solver="ct")
    M.init_turb_field(s)
    M.ct_faces_from_cells(s)
    for _ in range(3):
        dt, _ = M.compute_dt(s, 0.3)
        M.step(s, dt)
        _, folded = M.compute_dt(s, 0.3)
        s.device_ptr(0)  # invalidates the kept speed: the next call reduces the fields again
        _, fresh = M.compute_dt(s, 0.3)
        assert abs(folded - fresh) <= tol * fresh
    # a field-less workspace's speed of other arrays must not stand in for the state's
    g = M.allocate_state((16, 16, 16), 1 / 16, dtype=dtype)
    g.set_conserved(smooth_state((16, 16, 16), seed=3))
    dt, sm = M.compute_dt(g, 0.3)
    M.step(g, dt, ch=sm, integrator="cube")
    _, kept = M.compute_dt(g, 0.3)
    other = M.allocate_state((16, 16, 16), 1 / 16, dtype=dtype)
    other.set_conserved(2.0 * smooth_state((16, 16, 16), seed=4))
    M.max_wavespeed_of(g, [other.device_ptr(v) for v in range(9)])
    _, again = M.compute_dt(g, 0.3)
    assert abs(again - kept) <= tol * kept


def test_ct_float32_divergence_preserved_and_faces_from_cells():
    N = 16
    s = M.allocate_state((N, N, N), 1.0 / N, dtype=np.float32, riemann="hlld", mhd_solver="ct")
    M.init_turb_field(s)
    info = M.ct_faces_from_cells(s)
    assert info["max_divb_face_dx_over_brms"] < 1e-5
    assert info["cell_b_rel_l2_change"] < 1e-6
    divs = []
    M.evolve(s, 0.1, cfl=0.3, callback=lambda st, t, n: divs.append(M.max_divb_face(st) / N))
    assert max(divs) < 1e-5 and len(divs) > 5


@pytest.mark.parametrize("recon", ["plm", "plm_minmod", "ppm"])
@pytest.mark.parametrize("riemann", ["hlld", "hll", "llf"])
def test_ct_fused_kernel_matches_oracle_and_per_cell_kernel_float64(recon, riemann):
    """CT states step on the fused z-marching kernel (the :ref algorithm with every face
    solved once): against the float64 oracle (cells and
    faces) and against the per-cell kernel (every CT slot, the workspace included). 16^3
    has a ragged last tile row and four z chunks."""
    from test_oracle import ct_test_hydro, ct_test_potential
    N = 16
    kw = dict(riemann=riemann, recon=recon, mhd_solver="ct", smallr=F.FK_SMALLR, pfl=F.FK_SMALLP)
    g = M.allocate_state((N, N, N), 1.0 / N, dtype=np.float64, **kw)
    p = M.allocate_state((N, N, N), 1.0 / N, dtype=np.float64, **kw)
    M.set_ct_kernel(p, True)
    o = O.OracleState((N, N, N), 1.0 / N, **kw)
    U = ct_test_hydro(N, G)
    A = ct_test_potential(N)
    for st in (g, p):
        st.set_conserved(U)
        M.ct_faces_from_potential(st, *A)
    o.set_conserved(U)
    o.ct_faces_from_potential(*A)
    dt, _ = o.compute_dt(0.3)
    for _ in range(3):
        o.step(dt)
        M.step(g, dt)
        M.step(p, dt)
    Ug, Up = g.conserved(), p.conserved()
    for v in range(9):
        scale = max(np.abs(o.U[v]).max(), 1e-6)
        assert np.max(np.abs(Ug[v] - o.U[v])) <= 1e-12 * scale, v
        assert np.max(np.abs(Ug[v] - Up[v])) <= 1e-12 * scale, v
    for d in range(3):
        assert np.max(np.abs(g.ct_slot(d) - o.ct[d])) <= 1e-12 * np.abs(o.ct[d]).max()
    for q in range(21):
        a, b = g.ct_slot(q), p.ct_slot(q)
        assert np.max(np.abs(a - b)) <= 1e-12 * max(np.abs(b).max(), 1e-300), q
    assert M.max_divb_face(g) / N < 1e-12
    assert np.all(Ug[8] == 0)
    # the cell-field pass folded the finished state's CFL speed (timestep.jl:8-12)
    _, sm = M.compute_dt(g, 0.3)
    assert abs(sm - o.max_wavespeed()) <= 1e-14 * o.max_wavespeed()


def test_ct_fused_kernel_float32_matches_per_cell_kernel():
    """256^3-recipe turbulence at 32^3 (float32): the fused CT step against the per-cell
    one after 5 CFL steps, and the face divergence stays at rounding level."""
    N = 32
    s = M.allocate_state((N, N, N), 1.0 / N, dtype=np.float32, recon="plm_minmod", riemann="hlld", mhd_solver="ct")
    p = M.allocate_state((N, N, N), 1.0 / N, dtype=np.float32, recon="plm_minmod", riemann="hlld", mhd_solver="ct")
    M.set_ct_kernel(p, True)
    U = P.config_state("turb_256", seed=2, n=N).astype(np.float32)
    for st in (s, p):
        st.set_conserved(U)
        M.ct_faces_from_cells(st)
    for _ in range(5):
        dt, _ = M.compute_dt(s, 0.3)
        M.step(s, float(dt))
        M.step(p, float(dt))
    a, b = s.conserved().astype(np.float64), p.conserved().astype(np.float64)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6
    assert M.max_divb_face(s) / N < 1e-5


def test_errors_carry_reference_text():
    s = M.allocate_state((8, 8, 8), 1 / 8)
    with pytest.raises(M.MHDError, match="needs the GLM cleaning speed"):
        M.step(s, 1e-3)
    with pytest.raises(M.MHDError, match="unknown integrator :rk4"):
        M.step(s, 1e-3, ch=1.0, integrator="rk4")
    c = M.allocate_state((8, 8, 8), 1 / 8, mhd_solver="ct")
    with pytest.raises(M.MHDError, match="CT is implemented on the :ref integrator only"):
        M.step(c, 1e-3, integrator="cube")
    with pytest.raises(M.MHDError, match="max_divb_face needs a CT state"):
        M.max_divb_face(s)


# --- the fused z-marching kernel (integrator "cube") ------------------------------------
@pytest.mark.parametrize("dims", [(32, 16, 12), (16, 16, 5), (48, 32, 3), (16, 32, 40)])
@pytest.mark.parametrize("recon", ["plm", "ppm", "plm_minmod"])
def test_fused_float64_matches_oracle(dims, recon):
    U = smooth_state(dims, seed=11)
    g, o = pair(dims, U, recon=recon)
    dt, smax = o.compute_dt(0.3)
    for _ in range(2):
        o.step(dt, ch=smax)
        M.step(g, dt, ch=smax, integrator="cube")
    Ug = g.conserved()
    for v in range(9):
        assert np.max(np.abs(Ug[v] - o.U[v])) <= 1e-12 * max(np.abs(o.U[v]).max(), 1e-3), (v, dims, recon)
    # the fused epilogue's CFL maximum of the new state
    _, sm_dev = M.compute_dt(g, 0.3)
    assert abs(sm_dev - o.max_wavespeed()) <= 1e-14 * o.max_wavespeed()


@pytest.mark.parametrize("solver", ["hll", "llf"])
def test_fused_float64_other_solvers(solver):
    dims = (32, 16, 8)
    U = smooth_state(dims, seed=12)
    g, o = pair(dims, U, riemann=solver)
    dt, smax = o.compute_dt(0.3)
    o.step(dt, ch=smax)
    M.step(g, dt, ch=smax, integrator="cube")
    Ug = g.conserved()
    for v in range(9):
        assert np.max(np.abs(Ug[v] - o.U[v])) <= 1e-12 * max(np.abs(o.U[v]).max(), 1e-3), v


def test_fused_float32_config1_gates():
    n = 32
    U = P.config_state("bstatic_32", seed=1)
    g, o = pair((n, n, n), U, dtype=np.float32, recon="plm_minmod")
    r = M.allocate_state((n, n, n), 1.0 / n, dtype=np.float32, recon="plm_minmod")
    r.set_conserved(U)
    dt, smax = o.compute_dt(0.2)
    o.step(float(dt), ch=smax)
    M.step(g, float(dt), ch=smax, integrator="cube")
    M.step(r, float(dt), ch=smax, integrator="ref")
    Ug, Ur = g.conserved(), r.conserved()
    for v in range(9):
        assert relL2(Ug[v], o.U[v]) <= 1e-5, v
        assert relL2(Ug[v], Ur[v]) <= 2e-6, v
    # 49 more CFL steps against the float64 oracle
    o64 = O.OracleState((n, n, n), 1.0 / n, dtype=np.float64, recon="plm_minmod")
    o64.set_conserved(U)
    o64.step(float(dt), ch=smax)
    for _ in range(49):
        d64, s64 = o64.compute_dt(0.2)
        o64.step(float(d64), ch=s64)
        dg, sg = M.compute_dt(g, 0.2)
        M.step(g, float(dg), ch=sg, integrator="cube")
    Ug = g.conserved()
    errs = [relL2(Ug[v], o64.U[v]) for v in range(9)]
    assert max(errs) <= 1e-5, errs


@pytest.mark.parametrize("recon", ["plm_minmod", "ppm"])
def test_fused_float32_wide_offsets_equal_narrow_offsets(recon, monkeypatch):
    """Float fields below 4 GiB run the kernels with 32-bit byte offsets; 1024^3 and larger
    grids run the 64-bit-offset instances of the same code. Forced onto a small grid, the
    wide instances give the narrow ones' results bit for bit, the CFL fold included."""
    dims = (48, 26, 20)
    U = smooth_state(dims, seed=21)
    out = []
    for wide in ("0", "1"):
        monkeypatch.setenv("PMFMHD_WIDE_OFFSETS", wide)
        s = M.allocate_state(dims, 1.0 / 48, dtype=np.float32, recon=recon)
        s.set_conserved(U)
        dt, sm = M.compute_dt(s, 0.3)
        for _ in range(3):
            M.step(s, float(dt), ch=float(sm), integrator="cube")
        out.append((s.conserved(), M.compute_dt(s, 0.3)))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


@pytest.mark.parametrize("recon,riemann", [("plm_minmod", "hlld"), ("ppm", "hll")])
def test_ct_fused_float32_wide_offsets_equal_narrow_offsets(recon, riemann, monkeypatch):
    """The fused CT kernel's 32-bit-offset instance (fields and workspace slots as scalar
    bases plus 32-bit byte offsets) against its 64-bit-offset instance: cells and all 21
    CT slots bit for bit."""
    from test_oracle import ct_test_hydro, ct_test_potential
    n = 24
    out = []
    for wide in ("0", "1"):
        monkeypatch.setenv("PMFMHD_WIDE_OFFSETS", wide)
        s = M.allocate_state((n,) * 3, 1.0 / n, dtype=np.float32, recon=recon, riemann=riemann, mhd_solver="ct")
        s.set_conserved(ct_test_hydro(n, G))
        M.ct_faces_from_potential(s, *ct_test_potential(n))
        for _ in range(3):
            dt, _ = M.compute_dt(s, 0.3)
            M.step(s, float(dt))
        out.append((s.conserved(), [s.ct_slot(q) for q in range(21)]))
    assert np.array_equal(out[0][0], out[1][0])
    for q in range(21):
        assert np.array_equal(out[0][1][q], out[1][1][q]), q


@pytest.mark.parametrize("dims", [(24, 24, 24), (40, 30, 20), (33, 13, 17), (16, 12, 5), (64, 64, 64), (8, 5, 3),
                                  (17, 25, 9)])
@pytest.mark.parametrize("recon,riemann", [("plm_minmod", "hlld"), ("plm", "hll"), ("plm", "hlld")])
def test_ct_step_forming_edges_in_kernel_equals_the_workspace_path(dims, recon, riemann):
    """The fused float CT step that forms the edge EMFs itself (the default) against the
    one that writes workspace slots 3-14 and runs the separate edge pass (the reference's
    layout, pmf_set_ct_workspace): after four CFL steps the cells, the faces, the edge
    EMFs and the remainders agree bit for bit, on grids with ragged tiles and chunks and
    one smaller than a tile; slots 3-14 are left alone by the default path."""
    nx, ny, nz = dims
    rng = np.random.default_rng(11)
    U = np.zeros((9, nz, ny, nx))
    U[0] = 1.0 + 0.2 * rng.random((nz, ny, nx))
    for d in range(3):
        U[1 + d] = U[0] * 0.4 * rng.standard_normal((nz, ny, nx))
        U[5 + d] = 0.6 * rng.standard_normal((nz, ny, nx))
    U[4] = 1.5 + 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / U[0] + 0.5 * (U[5] ** 2 + U[6] ** 2 + U[7] ** 2)
    out = []
    for fill in (True, False):
        s = M.allocate_state(dims, 1.0 / nx, dtype=np.float32, recon=recon, riemann=riemann, mhd_solver="ct")
        M.set_ct_workspace(s, fill)
        s.set_conserved(U.reshape(9, -1).astype(np.float32))
        M.ct_faces_from_cells(s)
        slots0 = [s.ct_slot(q) for q in range(3, 15)]
        for _ in range(4):
            dt, _ = M.compute_dt(s, 0.3)
            M.step(s, float(dt))
        out.append((s.conserved(), [s.ct_slot(q) for q in range(21)], slots0))
    (Uw, Sw, _), (Ue, Se, S0) = out
    assert np.all(np.isfinite(Ue))
    assert np.array_equal(Uw, Ue)
    for q in (0, 1, 2, 15, 16, 17, 18, 19, 20):
        assert np.array_equal(Sw[q], Se[q]), q
    for q in range(3, 15):
        assert np.array_equal(Se[q], S0[q - 3]), q
    assert np.any(Sw[3] != S0[0])


def test_fused_rejects_a_one_cell_reflecting_axis():
    b = M.allocate_state((16, 16, 1), 1 / 16, bcs=("periodic", "periodic", "reflecting"))
    with pytest.raises(M.MHDError, match="reflecting axes two"):
        M.step(b, 1e-3, ch=1.0, integrator="cube")


@pytest.mark.parametrize("dims,bcs", [((20, 9, 6), ("periodic",) * 3), ((24, 16, 8), ("periodic",) * 3),
                                      ((8, 5, 3), ("outflow", "reflecting", "periodic")),
                                      ((40, 1, 1), ("outflow", "periodic", "periodic")),
                                      ((24, 24, 1), ("periodic",) * 3), ((36, 2, 3), ("reflecting",) * 3)])
@pytest.mark.parametrize("recon", ["plm", "ppm"])
def test_fused_any_grid_shape_float64_matches_oracle(dims, bcs, recon):
    """The fused kernel on grids whose nx is not a multiple of the 16-cell tile (a ragged
    last tile column, computed and not stored), smaller than one tile, or one cell thick
    in y and z (1-D and 2-D problems): equal to the per-cell step to round-off."""
    g, o = _bc_pair(dims, bcs, recon=recon)
    dt, smax = o.compute_dt(0.3)
    for _ in range(2):
        o.step(dt, ch=smax)
        M.step(g, dt, ch=smax, integrator="cube")
    Ug = g.conserved()
    for v in range(9):
        assert np.max(np.abs(Ug[v] - o.U[v])) <= 1e-12 * max(np.abs(o.U[v]).max(), 1e-3), (v, dims, bcs)
    _, sm_dev = M.compute_dt(g, 0.3)
    assert abs(sm_dev - o.max_wavespeed()) <= 1e-14 * o.max_wavespeed()


def _bc_pair(dims, bcs, dtype=np.float64, seed=13, **kw):
    U = smooth_state(dims, seed=seed)
    g = M.allocate_state(dims, 1.0 / dims[0], dtype=dtype, bcs=bcs, **kw)
    g.set_conserved(U)
    o = O.OracleState(dims, 1.0 / dims[0], dtype=dtype, bcs=bcs, **kw)
    o.set_conserved(U)
    return g, o


@pytest.mark.parametrize("bcs", [("outflow",) * 3, ("reflecting",) * 3, ("outflow", "reflecting", "periodic"),
                                 ("reflecting", "periodic", "outflow")])
@pytest.mark.parametrize("recon", ["plm", "ppm", "plm_minmod"])
def test_fused_boundaries_float64_match_oracle(bcs, recon):
    """step_cube! synthesises outflow / reflecting ghost cells when it loads its halo
    (integrator_cube.jl:318-320, boundary.jl:17-41), so the fused kernel runs on such grids
    and equals the per-cell step to round-off. 32x14x10: a ragged last tile row (14 rows
    in tiles of 12) and three z chunks (4, 4, 2) whose first and last meet the z
    boundaries."""
    g, o = _bc_pair((32, 14, 10), bcs, recon=recon)
    dt, smax = o.compute_dt(0.3)
    for _ in range(2):
        o.step(dt, ch=smax)
        M.step(g, dt, ch=smax, integrator="cube")
    Ug = g.conserved()
    for v in range(9):
        assert np.max(np.abs(Ug[v] - o.U[v])) <= 1e-12 * max(np.abs(o.U[v]).max(), 1e-3), (v, bcs, recon)
    _, sm_dev = M.compute_dt(g, 0.3)
    assert abs(sm_dev - o.max_wavespeed()) <= 1e-14 * o.max_wavespeed()


@pytest.mark.parametrize("riemann", ["hll", "llf"])
def test_fused_boundaries_other_solvers_float64(riemann):
    g, o = _bc_pair((16, 16, 9), ("reflecting", "outflow", "reflecting"), riemann=riemann, recon="ppm")
    dt, smax = o.compute_dt(0.3)
    for _ in range(2):
        o.step(dt, ch=smax)
        M.step(g, dt, ch=smax, integrator="cube")
    Ug = g.conserved()
    for v in range(9):
        assert np.max(np.abs(Ug[v] - o.U[v])) <= 1e-12 * max(np.abs(o.U[v]).max(), 1e-3), (v, riemann)


def test_fused_boundaries_float32_match_float32_oracle_and_ref_kernel():
    """float32 on a reflecting / outflow grid: the fused step against the float32 oracle
    (<= 1e-5 relative L2 after one step, the module's bar) and against the per-cell kernel
    over three steps."""
    bcs = ("reflecting", "outflow", "reflecting")
    g, o = _bc_pair((48, 20, 12), bcs, dtype=np.float32, recon="plm_minmod")
    r = M.allocate_state((48, 20, 12), 1 / 48, dtype=np.float32, bcs=bcs, recon="plm_minmod")
    r.set_conserved(smooth_state((48, 20, 12), seed=13))
    dt, smax = o.compute_dt(0.3)
    o.step(dt, ch=smax)
    M.step(g, dt, ch=smax, integrator="cube")
    Ug = g.conserved()
    assert max(relL2(Ug[v], o.U[v]) for v in range(8)) <= 1e-5
    M.step(r, dt, ch=smax, integrator="ref")
    for _ in range(2):
        M.step(g, dt, ch=smax, integrator="cube")
        M.step(r, dt, ch=smax, integrator="ref")
    Ug, Ur = g.conserved(), r.conserved()
    assert max(relL2(Ug[v], Ur[v]) for v in range(8)) <= 1e-6


class _HipArrays:
    """Nine separately hipMalloc'ed device arrays (the runtime libpmfmhd.so uses)."""

    def __init__(self, host):
        import ctypes
        self.ct = ctypes
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []
        for a in host:
            p = ctypes.c_void_p()
            assert self.hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
            assert self.hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0
            self.ptrs.append(p.value)
        self.n = host[0].size

    def get(self):
        out = []
        for p in self.ptrs:
            a = np.empty(self.n, dtype=np.float32)
            assert self.hip.hipMemcpy(a.ctypes.data_as(self.ct.c_void_p), self.ct.c_void_p(p), a.nbytes, 2) == 0
            out.append(a)
        return np.stack(out)

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(self.ct.c_void_p(p))


@pytest.mark.parametrize("integrator", ["ref", "cube"])
def test_step_into_caller_owned_arrays_is_bitwise_the_state_step(integrator):
    """The drop-in seam: nine separately allocated device arrays (as the reference's
    MHDState holds them), stepped through a field-less workspace."""
    dims = (32, 24, 10)
    U = smooth_state(dims, seed=21).astype(np.float32)
    s = M.allocate_state(dims, 1.0 / dims[0], dtype=np.float32)
    s.set_conserved(U)
    dt, smax = M.compute_dt(s, 0.3)
    decay = float(np.exp(np.float32(-M.GLM_CR) * smax * dt / np.float32(1.0 / dims[0])))
    if integrator == "ref":
        M.step_ref(s, float(dt), float(smax), decay)
    else:
        M.step_cube(s, float(dt), float(smax), decay)
    want = s.conserved()
    ins = _HipArrays([U[v].copy() for v in range(9)])
    outs = _HipArrays([np.zeros_like(U[v]) for v in range(9)])
    try:
        ws = M.allocate_state(dims, 1.0 / dims[0], dtype=np.float32, workspace=True)
        assert float(M.max_wavespeed_of(ws, ins.ptrs)) == float(smax)
        M.step_into(ws, ins.ptrs, outs.ptrs, float(dt), float(smax), decay, integrator=integrator)
        ws.synchronize()
        assert np.array_equal(outs.get(), want)
        with pytest.raises(M.MHDError, match="workspace without fields"):
            M.compute_dt(ws)
    finally:
        ins.free()
        outs.free()


@pytest.mark.gpu
def test_fused_chunk_depth_and_autotune_leave_results_unchanged():
    """The z-chunk depth only regroups work: every depth gives the same bits, and the
    autotuner picks one of its candidates without touching the state."""
    import pmfmhd as M
    from pmfmhd import problems as P
    U = P.config_state("bstatic_32", seed=3, n=48)
    n = 48
    outs = []
    for zc in (0, 4, 7, 48, 200):
        s = M.allocate_state((n, n, n), 1 / n, dtype=np.float32, recon="plm_minmod")
        s.set_conserved(U)
        M.set_chunk(s, zc)
        for _ in range(3):
            dt, sm = M.compute_dt(s, 0.2)
            M.step(s, float(dt), ch=float(sm), integrator="cube")
        outs.append(s.conserved())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    s = M.allocate_state((n, n, n), 1 / n, dtype=np.float32, recon="plm_minmod")
    s.set_conserved(U)
    before = s.conserved()
    zc, ms = M.autotune(s, reps=2)
    assert zc in (4, 8, 12, 16, 24, 32, 48) and ms > 0
    assert np.array_equal(s.conserved(), before)
    for _ in range(3):
        dt, sm = M.compute_dt(s, 0.2)
        M.step(s, float(dt), ch=float(sm), integrator="cube")
    assert np.array_equal(s.conserved(), outs[0])
    with pytest.raises(M.MHDError, match="at least 4"):
        M.set_chunk(s, 2)


def test_ct_chunk_depth_and_autotune_leave_results_unchanged():
    """The same for the fused CT step: the tuner writes only scratch fields and per-step
    workspace, so cells and faces are untouched, and every depth gives the same bits."""
    n = 48
    U = P.config_state("turb_256", seed=4, n=n).astype(np.float32)

    def ct_state(zc):
        s = M.allocate_state((n, n, n), 1 / n, dtype=np.float32, recon="plm_minmod", mhd_solver="ct")
        s.set_conserved(U)
        M.ct_faces_from_cells(s)
        M.set_chunk(s, zc)
        return s

    a = ct_state(0)
    before = (a.conserved(), [a.ct_slot(q) for q in (0, 1, 2, 18, 19, 20)])
    zc, ms = M.autotune(a, reps=2)
    assert zc in (4, 8, 12, 16, 24, 32, 48) and ms > 0
    assert np.array_equal(a.conserved(), before[0])
    for q, b in zip((0, 1, 2, 18, 19, 20), before[1]):
        assert np.array_equal(a.ct_slot(q), b), q
    b = ct_state(7)
    for _ in range(3):
        dt, _ = M.compute_dt(a, 0.3)
        M.step(a, float(dt))
        M.step(b, float(dt))
    assert np.array_equal(a.conserved(), b.conserved())
    for q in range(3):
        assert np.array_equal(a.ct_slot(q), b.ct_slot(q))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("recon", ["plm", "plm_minmod", "ppm"])
@pytest.mark.parametrize("riemann", ["hlld", "hll", "llf"])
def test_uniform_state_is_a_fixed_point_bit_for_bit(dtype, recon, riemann):
    """Every face of a uniform periodic state has the same two states, so every flux
    difference must cancel exactly: one step on either integrator, a CT step, and the
    drag step with a zero target velocity change nothing but the drag's own velocity
    relaxation (here the state is at rest in the target frame, so nothing at all)."""
    n = 16
    rho, v, p, B = 1.3, (0.4, -0.2, 0.3), 0.8, (0.5, 0.2, -0.1)
    e = p / (G - 1) + 0.5 * rho * sum(x * x for x in v) + 0.5 * sum(b * b for b in B)
    U = np.zeros((9, n ** 3), dtype)
    for k, val in enumerate((rho, rho * v[0], rho * v[1], rho * v[2], e, B[0], B[1], B[2], 0.0)):
        U[k] = dtype(val)
    for integrator in ("ref", "cube"):
        s = M.allocate_state((n, n, n), 1.0 / n, dtype=dtype, recon=recon, riemann=riemann)
        s.set_conserved(U)
        dt, sm = M.compute_dt(s, 0.3)
        M.step(s, float(dt), ch=float(sm), integrator=integrator)
        assert np.array_equal(s.conserved(), U), integrator
    c = M.allocate_state((n, n, n), 1.0 / n, dtype=dtype, recon=recon, riemann=riemann, mhd_solver="ct")
    c.set_conserved(U)
    M.ct_faces_from_cells(c)
    Uc = c.conserved()
    dt, _ = M.compute_dt(c, 0.3)
    M.step(c, float(dt))
    assert np.array_equal(c.conserved(), Uc)
    r = U.copy()
    r[1:4] = 0
    r[4] = dtype(p / (G - 1) + 0.5 * sum(b * b for b in B))
    d = M.allocate_state((n, n, n), 1.0 / n, dtype=dtype, recon=recon, riemann=riemann)
    d.set_conserved(r)
    dt, sm = M.compute_dt(d, 0.3)
    M.step_drag(d, float(dt), float(sm), 0.99, 0.3, integrator="cube")
    assert np.array_equal(d.conserved(), r)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("fac,bcs", [(1.0, ("periodic",) * 3), (1.4, ("outflow", "reflecting", "periodic"))])
def test_advance_with_device_dt_equals_host_driven_steps(dtype, fac, bcs):
    """pmf_advance computes each dt, ch and psi decay on the device from the speed the
    previous step folded (evolve!'s rule, driver.jl:48-71); the same steps driven from the
    host (compute_dt, then step with ch = glm_ch_fac * smax and dt scaled by smax/max(smax,
    ch), as evolve! does) give the same state: bit for bit in float32, to rounding of the
    psi factor's exp in float64. The elapsed time is the sum of the host's dts."""
    dims = (48, 40, 32)
    U = smooth_state(dims, seed=31)
    kw = dict(dtype=dtype, recon="plm_minmod", bcs=bcs)
    a = M.allocate_state(dims, 1.0 / dims[0], **kw)
    b = M.allocate_state(dims, 1.0 / dims[0], **kw)
    a.set_conserved(U)
    b.set_conserved(U)
    t_host = 0.0
    for _ in range(6):
        dt, sm = M.compute_dt(a, 0.3)
        if dtype == np.float32:
            smf = np.float32(sm)
            chf = np.float32(fac) * smf
            ch = float(chf)
            dtf = float(dt) * float(smf) / float(max(smf, chf))
        else:
            ch = fac * float(sm)
            dtf = float(dt) * float(sm) / max(float(sm), ch)
        M.step(a, dtf, ch=ch, integrator="cube")
        t_host += dtf
    t, last = M.advance(b, 6, cfl=0.3, glm_ch_fac=fac)
    assert t == t_host and last == dtf
    Ua, Ub = a.conserved(), b.conserved()
    if dtype == np.float32:
        assert np.array_equal(Ua, Ub)
    else:
        scale = np.abs(Ua).max(axis=1, keepdims=True)
        assert np.max(np.abs(Ua - Ub) / scale) < 1e-14
    assert M.compute_dt(b, 0.3)[1] == M.compute_dt(a, 0.3)[1]
    with pytest.raises(M.MHDError, match="constrained-transport states step with pmf_step"):
        M.advance(M.allocate_state((16,) * 3, 1 / 16, dtype=dtype, mhd_solver="ct"), 1)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("recon", ["plm_minmod", "plm"])
@pytest.mark.parametrize("riemann", ["hlld", "hll", "llf"])
@pytest.mark.parametrize("dims,bcs", [((40, 30, 20), None), ((33, 26, 17), ("outflow", "reflecting", "periodic"))])
def test_fused_step_equals_the_per_cell_step_bitwise(dtype, recon, riemann, dims, bcs):
    """The fused z-marching step and the per-cell step (the reference's step_ref! layout,
    integrator_ref.jl:162-220) share the PLM Hancock half-step, its offsets and the face
    solvers, so two CFL steps of the turbulence recipe give the same bits in every cell:
    ragged tiles and chunks, and outflow / reflecting axes included."""
    n = max(dims)
    U = P.config_state("turb_256", seed=2, n=n).reshape(9, n, n, n)[:, :dims[2], :dims[1], :dims[0]]
    U = np.ascontiguousarray(U.reshape(9, -1))
    out = []
    for integrator in ("cube", "ref"):
        s = M.allocate_state(dims, 1.0 / n, dtype=dtype, recon=recon, riemann=riemann, bcs=bcs)
        s.set_conserved(U)
        for _ in range(2):
            dt, sm = M.compute_dt(s, 0.2)
            M.step(s, float(dt), ch=float(sm), integrator=integrator)
        out.append(s.conserved())
    assert np.array_equal(out[0], out[1])
