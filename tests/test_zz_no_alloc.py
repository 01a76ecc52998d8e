"""pmf_step_into_static, the step that allocates nothing (its workspace is a library-held
device array), through the C ABI against pmf_step_into on the same inputs: two calls in a
row, the second on a smaller grid, for the fused GLM step, the per-cell GLM step and the
CT step, in float32 and float64. Both run the same kernels, so the outputs are equal bit
for bit; the CFL speed the static call leaves in the caller's device double is the
output's max_wavespeed!."""
import ctypes

import numpy as np
import pytest

import pmfmhd as M
from pmfmhd._lib import lib
from test_gpu_parity import smooth_state
from test_oracle import ct_test_hydro, ct_test_potential

pytestmark = pytest.mark.gpu
G = 5.0 / 3.0


def _buffers(arrays):
    out = []
    for a in arrays:
        a = np.ascontiguousarray(a)
        b = M.DeviceBuffer(a.nbytes)
        b.upload(a)
        out.append(b)
    return out


def _ptrs(bufs):
    return (ctypes.c_void_p * 9)(*[b.ptr for b in bufs])


def _inputs(solver, dims, dtype):
    """Conserved fields (and, for CT, the 21 packed slots) of a state set up on the GPU."""
    if solver == "ct":
        n = dims[0]
        s = M.allocate_state(dims, 1.0 / n, dtype=dtype, recon="plm", riemann="hll", mhd_solver="ct")
        s.set_conserved(ct_test_hydro(n, G))
        M.ct_faces_from_potential(s, *ct_test_potential(n))
        return s, s.conserved(), np.concatenate([s.ct_slot(q) for q in range(21)])
    s = M.allocate_state(dims, 1.0 / dims[0], dtype=dtype, recon="ppm", riemann="hlld")
    s.set_conserved(smooth_state(dims, seed=sum(dims)).astype(dtype))
    return s, s.conserved(), None


@pytest.mark.parametrize("solver,integrator", [("glm", "cube"), ("glm", "ref"), ("ct", "ref")])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_static_step_equals_step_into_on_two_calls(solver, integrator, dtype):
    L = lib()
    icode = 1 if integrator == "cube" else 0
    shapes = ((24, 24, 24), (16, 16, 16)) if solver == "ct" else ((40, 36, 28), (24, 20, 12))
    for dims in shapes:  # the second call is smaller and finds the previous call's workspace contents
        s, U, slots = _inputs(solver, dims, dtype)
        dt, smax = M.compute_dt(s, 0.3)
        ch = 0.0 if solver == "ct" else float(smax)
        decay = 1.0 if solver == "ct" else 0.95
        nc = s.ncells
        ins = _buffers(list(U))
        out_a = _buffers([np.zeros(nc, dtype)] * 9)
        out_b = _buffers([np.zeros(nc, dtype)] * 9)
        ct_a = _buffers([slots])[0] if slots is not None else None
        ct_b = _buffers([slots])[0] if slots is not None else None
        smax_dev = M.DeviceBuffer(8)
        ws = M.allocate_state(dims, 1.0 / dims[0], dtype=dtype, recon=s.recon, riemann=s.riemann,
                              mhd_solver=solver, workspace=True)
        M.step_into(ws, [b.ptr for b in ins], [b.ptr for b in out_a], float(dt), ch, decay, integrator=integrator,
                    ct_ptr=None if ct_a is None else ct_a.ptr)
        ws.synchronize()
        rc = L.pmf_step_into_static(ctypes.byref(s._p), _ptrs(ins), _ptrs(out_b),
                                    None if ct_b is None else ctypes.c_void_p(ct_b.ptr), float(dt), ch, decay, icode,
                                    None, ctypes.c_void_p(smax_dev.ptr))
        assert rc == 0, L.pmf_last_error().decode()
        got = np.stack([b.download(dtype, nc) for b in out_b])
        want = np.stack([b.download(dtype, nc) for b in out_a])
        assert np.array_equal(got, want), dims
        if slots is not None:
            assert np.array_equal(ct_b.download(dtype, 21 * nc), ct_a.download(dtype, 21 * nc)), dims
        fresh = float(M.max_wavespeed_of(ws, [b.ptr for b in out_a]))
        folded = float(smax_dev.download(np.float64, 1)[0])
        # the fused GLM epilogue and CT's cell-field pass fold the speed in their own kernels
        # (FMA contraction may differ from the separate reduction); the per-cell GLM step
        # runs the separate reduction itself
        tol = 0.0 if (solver == "glm" and integrator == "ref") else (4e-7 if dtype == np.float32 else 1e-15)
        assert abs(folded - fresh) <= tol * fresh, (dims, folded, fresh)


def test_static_step_refuses_a_grid_beyond_its_workspace():
    """A grid whose fused step needs more partial slots than the static workspace holds
    is refused before anything runs (no field is touched: the pointers are never read)."""
    L = lib()
    p = M._lib.pmf_params()
    assert L.pmf_params_default(ctypes.byref(p), 8192, 8192, 64, 1.0 / 8192) == 0
    bufs = _buffers([np.zeros(16, np.float32)] * 9)
    rc = L.pmf_step_into_static(ctypes.byref(p), _ptrs(bufs), _ptrs(bufs), None, 1e-3, 1.0, 1.0, 1, None, None)
    assert rc != 0
    assert "static workspace" in L.pmf_last_error().decode() and "pmf_step_into" in L.pmf_last_error().decode()
