"""The batched ensemble: every cube must come out exactly as a separate state stepped
with compute_dt + step(integrator="cube") would (the per-cube dt stage on the device
repeats the host's float32 arithmetic), whatever the graph replay, the tile chunking
or the persistent grid size."""
import ctypes

import numpy as np
import pytest

from pmfmhd import problems as P


def _members(edges):
    out = []
    for c, n in enumerate(edges):
        u = P.hash_uniform(c, 21, np.arange(2, dtype=np.int64))
        out.append(dict(cube=c, n=n, seed=c, brms=0.1 + 1.9 * float(u[0]), b_index=-2.5 + 4.5 * float(u[1])))
    return out


def _single_run(U, n, steps, cfl=0.2, recon="plm_minmod", riemann="hlld"):
    import pmfmhd as M
    s = M.allocate_state((n, n, n), 1 / n, dtype=np.float32, recon=recon, riemann=riemann)
    s.set_conserved(U)
    t = 0.0
    for _ in range(steps):
        dt, smax = M.compute_dt(s, cfl)
        ch = float(np.float32(smax))
        dtf = float(dt) * float(np.float32(smax)) / max(float(np.float32(smax)), ch)
        M.step(s, dtf, ch=ch, integrator="cube")
        t += dtf
    return s.conserved(), t


def test_ensemble_rejects_bad_shapes_before_touching_the_gpu():
    from pmfmhd._lib import MHDError
    from pmfmhd.ensemble import Ensemble
    with pytest.raises(MHDError, match="multiples of 16"):
        Ensemble([32, 24])
    with pytest.raises(MHDError, match="at least one cube"):
        Ensemble([])


def test_ensemble_layout_and_member_draws():
    m = P.ensemble_members()
    assert [x["n"] for x in m] == [32] * 32 + [64] * 24 + [128] * 8
    cells = sum(x["n"] ** 3 for x in m)
    assert cells == 24_117_248
    b = np.array([x["brms"] for x in m])
    nb = np.array([x["b_index"] for x in m])
    assert b.min() >= 0.1 and b.max() <= 2.0 and nb.min() >= -2.5 and nb.max() <= 2.0
    assert len(set(np.round(b, 12))) == 64  # independent draws per cube
    U = P.ensemble_state(m[0])
    # rms |B| as drawn
    Bc = U[5:8]
    assert np.sqrt((Bc ** 2).sum(axis=0).mean()) == pytest.approx(m[0]["brms"], rel=1e-12)


@pytest.mark.gpu
def test_ensemble_matches_single_cube_runs_graph_and_eager():
    from pmfmhd.ensemble import Ensemble
    edges = [32, 48, 16, 64, 32, 16]
    mem = _members(edges)
    Us = [P.ensemble_state(m) for m in mem]
    steps = 7  # three graph replays and one eager step
    outs = {}
    for graph in (True, False):
        e = Ensemble(edges)
        for c, U in enumerate(Us):
            e.upload(c, U)
        e.run(steps, graph=graph)
        outs[graph] = ([e.download(c) for c in range(len(edges))], e.clock())
    for c, n in enumerate(edges):
        ref, t = _single_run(Us[c], n, steps)
        got = outs[True][0][c]
        assert np.array_equal(got, outs[False][0][c]), f"cube {c}: graph and eager differ"
        d = np.abs(got.astype(np.float64) - ref).max() / np.abs(ref).max()
        print(f"cube {c} (n={n}): max |ensemble - single| / max = {d:.3e}, t {outs[True][1][0][c]!r} vs {t!r}")
        assert d == 0.0, f"cube {c}: {d}"
        assert outs[True][1][0][c] == t


@pytest.mark.gpu
@pytest.mark.parametrize("recon,riemann", [("ppm", "hlld"), ("plm", "hll"), ("plm_minmod", "llf")])
def test_ensemble_other_schemes_match_single_cube_runs(recon, riemann):
    """Every reconstruction and Riemann solver of the ensemble (its own kernel instances)
    against separate single-state runs, bit for bit."""
    from pmfmhd.ensemble import Ensemble
    edges = [32, 16, 48]
    mem = _members(edges)
    Us = [P.ensemble_state(m) for m in mem]
    e = Ensemble(edges, recon=recon, riemann=riemann)
    for c, U in enumerate(Us):
        e.upload(c, U)
    e.run(5)
    t_ens = e.clock()[0]
    for c, n in enumerate(edges):
        ref, t = _single_run(Us[c], n, 5, recon=recon, riemann=riemann)
        assert np.array_equal(e.download(c), ref), (c, recon, riemann)
        assert t_ens[c] == t


@pytest.mark.gpu
def test_ensemble_result_independent_of_chunking_grid_and_tuner():
    from pmfmhd.ensemble import Ensemble
    edges = [64, 32, 32, 48]
    Us = [P.ensemble_state(m) for m in _members(edges)]
    res = []
    for cfg in (None, ([4, 32, 8, 16], 256), ([64, 4, 32, 48], 1024), "tune"):
        e = Ensemble(edges)
        for c, U in enumerate(Us):
            e.upload(c, U)
        if cfg == "tune":
            before = [e.download(c) for c in range(len(edges))]
            ms = e.tune(reps=2)
            assert ms > 0
            zc, nwg, nit = e.config()
            print(f"tuned: zc {list(zc)} nwg {nwg} items {nit} ({ms:.3f} ms/step)")
            for c in range(len(edges)):
                assert np.array_equal(e.download(c), before[c])  # tuning leaves the state alone
            assert (e.clock()[0] == 0).all()
        elif cfg is not None:
            e.set_config(*cfg)
        e.run(4)
        res.append([e.download(c) for c in range(len(edges))])
    for r in res[1:]:
        for c in range(len(edges)):
            assert np.array_equal(r[c], res[0][c])


@pytest.mark.gpu
def test_full_ensemble_100_steps_cube_by_cube():
    """64 cubes (32 x 32^3, 24 x 64^3, 8 x 128^3), 100 steps each with its own dt,
    autotuned, against independent single-cube runs."""
    from pmfmhd.ensemble import Ensemble
    mem = P.ensemble_members()
    edges = [m["n"] for m in mem]
    e = Ensemble(edges)
    Us = []
    for c, m in enumerate(mem):
        U = P.ensemble_state(m).astype(np.float32)
        Us.append(U)
        e.upload(c, U)
    ms = e.tune(reps=2)
    sec = e.run(100)
    cells = sum(n ** 3 for n in edges)
    print(f"ensemble: tuned {ms:.3f} ms/step, 100 steps {sec * 1e3:.1f} ms = {cells * 100 / sec / 1e9:.2f} Gcell/s")
    t_ens = e.clock()[0]
    worst = 0.0
    for c, n in enumerate(edges):
        ref, t = _single_run(Us[c], n, 100)
        got = e.download(c)
        d = np.abs(got.astype(np.float64) - ref).max() / np.abs(ref).max()
        worst = max(worst, d)
        assert t_ens[c] == t, c
        assert np.isfinite(got).all()
    print(f"worst cube difference {worst:.3e}")
    assert worst == 0.0
