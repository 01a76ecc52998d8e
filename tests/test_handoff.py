"""The production runner's raw state handoff (src/production/runner.jl:3814-3883):
mhd_U_1.f32 ... mhd_U_9.f32, float32, x fastest. The reference's own "joint PMF handoff
import" case (test/inherited/test_pmf_checkpoint_helpers.jl:216-231: field i filled with i,
density 1.25, energy 8, at N = 4) and its validation errors, export -> import round trips in
both precisions, and the file layout against numpy's own raw float32 writer."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _write_raw(directory, arrays):
    os.makedirs(directory, exist_ok=True)
    for i, a in enumerate(arrays):
        np.asarray(a, dtype="<f4").tofile(os.path.join(directory, f"mhd_U_{i + 1}.f32"))


def test_reference_joint_handoff_case(tmp_path):
    import pmfmhd as M
    N = 4
    expected = [np.full(N ** 3, float(i), np.float32) for i in range(1, 10)]
    expected[0][:] = 1.25
    expected[4][:] = 8.0
    _write_raw(tmp_path, expected)
    s = M.allocate_state((N, N, N), 1.0 / N, dtype=np.float32, riemann="hll", recon="ppm")
    M.import_mhd_raw(s, tmp_path)
    for i in range(9):
        assert np.array_equal(s.download(i), expected[i])
    dt, smax = M.compute_dt(s, 0.3)  # the imported state is the one the next dt sees
    assert np.isfinite(dt) and smax > 0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_export_import_round_trip_and_layout(tmp_path, dtype):
    import pmfmhd as M
    from pmfmhd import problems as P
    dims = (12, 8, 6)
    U = P.config_state("bstatic_32", seed=5, n=None)[:, : np.prod(dims)].astype(dtype)
    a = M.allocate_state(dims, 1 / 12, dtype=dtype)
    a.set_conserved(U)
    M.export_mhd_raw(a, tmp_path)
    names = sorted(os.listdir(tmp_path))
    assert names == sorted(f"mhd_U_{i}.f32" for i in range(1, 10))  # no .tmp left behind
    for i in range(9):
        raw = np.fromfile(os.path.join(tmp_path, f"mhd_U_{i + 1}.f32"), dtype="<f4")
        assert raw.size == np.prod(dims)
        assert np.array_equal(raw, U[i].astype(np.float32))  # x fastest, float32 on disk
    b = M.allocate_state(dims, 1 / 12, dtype=dtype)
    M.import_mhd_raw(b, tmp_path)
    assert np.array_equal(b.conserved(), U.astype(np.float32).astype(dtype))


def test_validation_errors_and_a_refused_import_leaves_the_state(tmp_path):
    import pmfmhd as M
    N = 4
    good = [np.full(N ** 3, 1.0 + i, np.float32) for i in range(9)]
    s = M.allocate_state((N, N, N), 1.0 / N, dtype=np.float32)
    s.set_conserved(np.stack(good) * 0 + 3.0)
    before = s.conserved()
    with pytest.raises(M.MHDError, match=r"bad raw Float32 size for .*mhd_U_1\.f32: got 0 bytes, expected 256"):
        M.import_mhd_raw(s, tmp_path / "missing")
    short = [g.copy() for g in good]
    short[6] = short[6][:-1]
    _write_raw(tmp_path / "short", short)
    with pytest.raises(M.MHDError, match=r"mhd_U_7\.f32: got 252 bytes, expected 256"):
        M.import_mhd_raw(s, tmp_path / "short")
    neg = [g.copy() for g in good]
    neg[0][5] = -0.5
    _write_raw(tmp_path / "neg", neg)
    with pytest.raises(M.MHDError, match="imported MHD state has invalid minimum density -0.5"):
        M.import_mhd_raw(s, tmp_path / "neg")
    bad_e = [g.copy() for g in good]
    bad_e[4][7] = np.nan
    _write_raw(tmp_path / "nan", bad_e)
    with pytest.raises(M.MHDError, match="imported MHD state has invalid minimum energy nan"):
        M.import_mhd_raw(s, tmp_path / "nan")
    assert np.array_equal(s.conserved(), before)
