"""Passive species transport (passive_transport.jl): the numpy restatement pinned on the
CPU, then the HIP kernels against it."""
import numpy as np
import pytest

import passive_np as PN


def _state(n, vx, vy, vz, rho=None, seed=0):
    rng = np.random.default_rng(seed)
    U = np.zeros((9, n ** 3))
    U[0] = rho if rho is not None else 1 + 0.2 * rng.random(n ** 3)
    U[1], U[2], U[3] = U[0] * vx, U[0] * vy, U[0] * vz
    U[4] = 2.0
    return U


def test_uniform_composition_is_preserved_exactly():
    n = 8
    rng = np.random.default_rng(1)
    U = _state(n, rng.normal(0, 1, n ** 3), rng.normal(0, 1, n ** 3), rng.normal(0, 1, n ** 3))
    a = np.full(n ** 3, 0.3)
    out = PN.backtrace(a, U, n, 0.7)
    np.testing.assert_array_equal(out, U[0] * 0.3)
    s = PN.backtrace(a, U, n, 0.7, mode=2)
    np.testing.assert_array_equal(s, a)


def test_zero_velocity_is_the_identity_and_one_cell_moves_are_shifts():
    n = 8
    rng = np.random.default_rng(2)
    a = rng.random(n ** 3)
    U = _state(n, 0.0, 0.0, 0.0)
    np.testing.assert_array_equal(PN.backtrace(a, U, n, 0.5, mode=2), a)
    # dtdx * v = (1, 0, -2) cells exactly: sample from (i-1, j, k+2)
    U = _state(n, 0.25, 0.0, -0.5, rho=np.ones(n ** 3))
    got = PN.backtrace(a, U, n, 4.0, mode=2).reshape(n, n, n)
    want = np.roll(np.roll(a.reshape(n, n, n), 1, axis=2), -2, axis=0)
    np.testing.assert_array_equal(got, want)


def test_linear_profiles_move_exactly_by_a_fractional_displacement():
    n = 16
    c = np.arange(n ** 3)
    i, j = c % n, (c // n) % n
    a = 0.01 * i + 0.02 * j  # linear away from the periodic seam
    U = _state(n, 0.3, -0.2, 0.0, rho=np.full(n ** 3, 2.0))
    got = PN.backtrace(a, U, n, 1.0, mode=2)
    want = 0.01 * (i - 0.3) + 0.02 * (j + 0.2)
    inner = (i >= 1) & (i <= n - 2) & (j >= 1) & (j <= n - 2)
    np.testing.assert_allclose(got[inner], want[inner], rtol=0, atol=1e-15)


def test_fraction_clamps_and_the_coupled_bound():
    rho = np.array([2.0, 2.0, 2.0, 1e-20])
    s1 = np.array([0.5, 3.0, -1.0, 1.0])
    np.testing.assert_allclose(PN.species_to_fractions(s1, rho), [0.25, 1.0, 0.0, 1.0])
    s2 = np.array([1.2, 0.4, 0.4, 0.0])
    q1, q2 = PN.species_to_fractions(s1, rho, 1.0, s2)
    np.testing.assert_allclose(q2, [0.6, 0.2, 0.2, 0.0])
    np.testing.assert_allclose(q1, [0.25, 0.8, 0.0, 1.0])
    n = 4
    U = _state(n, 0.0, 0.0, 0.0, rho=np.full(n ** 3, 3.0))
    o1, o2 = PN.backtrace(np.full(n ** 3, 0.9), U, n, 0.1, maximum_total=1.0, mode=1, a2=np.full(n ** 3, 0.5))
    np.testing.assert_allclose(o2, 1.5)
    np.testing.assert_allclose(o1, 1.5)  # 0.9 bounded by 1 - 0.5


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gpu_species_transport_matches_restatement(dtype):
    import pmfmhd as M
    n = 16
    rng = np.random.default_rng(3)
    U = _state(n, rng.normal(0, 0.5, n ** 3), rng.normal(0, 0.5, n ** 3), rng.normal(0, 0.5, n ** 3), seed=4)
    U = U.astype(dtype)
    s = M.allocate_state((n, n, n), 1 / n, dtype=dtype)
    s.set_conserved(U)
    dt = 0.03
    dtdx = dtype(dt) / dtype(1 / n) if dtype == np.float64 else np.float32(np.float32(dt) / np.float32(1 / n))
    tol = 1e-13 if dtype == np.float64 else 2e-6
    sp1 = (U[0] * rng.random(n ** 3) * 0.6).astype(dtype)
    sp2 = (U[0] * rng.random(n ** 3) * 0.3).astype(dtype)
    b1, b2 = M.DeviceBuffer(sp1.nbytes), M.DeviceBuffer(sp2.nbytes)
    t1, t2 = M.DeviceBuffer(sp1.nbytes), M.DeviceBuffer(sp2.nbytes)
    b1.upload(sp1)
    b2.upload(sp2)
    M.species_to_fractions(s, b1, b2, maximum_total=0.8)
    q1, q2 = PN.species_to_fractions(sp1, U[0], 0.8, sp2)
    np.testing.assert_allclose(b1.download(dtype, n ** 3), q1, rtol=tol, atol=tol)
    np.testing.assert_allclose(b2.download(dtype, n ** 3), q2, rtol=tol, atol=tol)
    M.advect_species(s, b1, t1, dt, species2=b2, scratch2=t2, maximum_total=0.8)
    o1, o2 = PN.backtrace(q1, U, n, dtdx, 0.8, mode=1, a2=q2)
    np.testing.assert_allclose(b1.download(dtype, n ** 3), o1, rtol=tol, atol=tol)
    np.testing.assert_allclose(b2.download(dtype, n ** 3), o2, rtol=tol, atol=tol)
    # single fraction and the unconstrained scalar
    a = rng.normal(0, 1, n ** 3).astype(dtype)
    b1.upload(a)
    M.advect_passive_scalar(s, b1, t1, dt)
    np.testing.assert_allclose(b1.download(dtype, n ** 3), PN.backtrace(a, U, n, dtdx, mode=2), rtol=tol, atol=tol)
    f = (0.5 + 0.3 * rng.random(n ** 3)).astype(dtype)
    b1.upload(f)
    M.advect_species(s, b1, t1, dt, maximum_total=0.7)
    np.testing.assert_allclose(b1.download(dtype, n ** 3), PN.backtrace(f, U, n, dtdx, 0.7), rtol=tol, atol=tol)
    c = M.allocate_state((16, 16, 8), 1 / 16, dtype=dtype)
    with pytest.raises(M.MHDError, match="species remap currently requires a cubic grid"):
        M.advect_species(c, b1, t1, dt)
