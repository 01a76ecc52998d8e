"""pmfmhd — the MHDKernels GLM-HLLD step on MI355X, from Python.

A host-side mirror of the lib/MHDKernels interface over the C ABI of libpmfmhd.so:
``allocate_state`` (state.jl:123-141), ``step`` (driver.jl:20-37), ``evolve``
(driver.jl:48-71), ``compute_dt`` / ``max_wavespeed`` (timestep.jl:17-41), ``max_divb`` /
``conserved_totals`` (diagnostics.jl), the CT helpers (ct.jl) and the problem
initialisers (problems.jl). Names, argument meaning and error text follow the reference
(Julia's ``step!`` is ``step`` here). All compute runs in the HIP library; there is no CPU
fallback, and a missing library raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import problems
from ._lib import MHDError, check, lib, pmf_params
from .ctinit import faces_from_cells

__all__ = ["GLM_CR", "MHDState", "MHDError", "allocate_state", "step", "step_ref", "step_cube", "evolve",
           "compute_dt", "max_wavespeed", "max_divb", "divb_norm", "conserved_totals", "max_divb_face",
           "ct_faces_from_cells", "ct_faces_from_potential", "riemann_flux", "fields_to_host",
           "init_alfven_wave", "init_orszag_tang", "init_brio_wu", "init_turb_field", "init_field_loop",
           "init_config", "problems", "step_into", "max_wavespeed_of"]

GLM_CR = 0.18  # driver.jl:8
_RIEMANN = {"hlld": 0, "hll": 1, "llf": 2}
_RECON = {"plm": 0, "ppm": 1, "plm_minmod": 2}
_BC = {"periodic": 0, "outflow": 1, "reflecting": 2}
_INTEG = {"ref": 0, "cube": 1}


def _code(table, kind, value, have):
    try:
        return table[value]
    except KeyError:
        raise MHDError(f"unknown {kind} :{value} (have {have})") from None


class MHDState:
    """Device state: 9 SoA conserved fields, scratch set and, for CT, the face buffer.
    With ``workspace=True`` it holds no fields (the caller owns its arrays; see step_into)."""

    def __init__(self, params: pmf_params, dtype, riemann, recon, bcs, solver, workspace=False):
        self._p = params
        self.dtype = np.dtype(dtype)
        self.riemann, self.recon, self.bcs, self._solver = riemann, recon, bcs, solver
        self.workspace = workspace
        h = ctypes.c_void_p()
        create = lib().pmf_create_workspace if workspace else lib().pmf_create
        check(create(ctypes.byref(params), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().pmf_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    @property
    def dims(self):
        return (self._p.nx, self._p.ny, self._p.nz)

    @property
    def ncells(self) -> int:
        return int(np.prod(self.dims))

    @property
    def dx(self):
        return self.dtype.type(self._p.dx)

    @property
    def gamma(self):
        return self.dtype.type(self._p.gamma)

    @property
    def uses_ct(self) -> bool:
        return self._solver == "ct"

    @property
    def mhd_solver(self) -> str:
        return self._solver

    def upload(self, var: int, a):
        a = np.ascontiguousarray(np.asarray(a).reshape(-1), dtype=self.dtype)
        check(lib().pmf_upload(self._h, var, a.ctypes.data_as(ctypes.c_void_p), a.size))

    def download(self, var: int) -> np.ndarray:
        a = np.empty(self.ncells, dtype=self.dtype)
        check(lib().pmf_download(self._h, var, a.ctypes.data_as(ctypes.c_void_p), a.size))
        return a

    def set_conserved(self, U):
        """Upload a (9, ncells) conserved state (cast to the state's precision)."""
        U = np.asarray(U)
        if U.shape[0] != 9:
            raise MHDError("conserved state must have 9 components")
        for v in range(9):
            self.upload(v, U[v])

    def fields_to_host(self):
        """state.jl:144: host copies of the 9 conserved fields."""
        return tuple(self.download(v) for v in range(9))

    def conserved(self) -> np.ndarray:
        return np.stack(self.fields_to_host())

    def ct_slot(self, q: int) -> np.ndarray:
        """Host copy of CT slot q (0..2 = bxf, byf, bzf; state.jl:17-37 order)."""
        a = np.empty(self.ncells, dtype=self.dtype)
        check(lib().pmf_ct_download(self._h, q, a.ctypes.data_as(ctypes.c_void_p), a.size))
        return a

    def device_ptr(self, var: int) -> int:
        p = ctypes.c_void_p()
        check(lib().pmf_device_ptr(self._h, var, ctypes.byref(p)))
        return p.value

    def scratch_ptr(self, var: int) -> int:
        p = ctypes.c_void_p()
        check(lib().pmf_scratch_ptr(self._h, var, ctypes.byref(p)))
        return p.value

    def ct_ptr(self, slot: int = 0) -> int:
        """Device address of CT slot q; slot 0 starts the 21-slot packed buffer."""
        p = ctypes.c_void_p()
        check(lib().pmf_ct_ptr(self._h, slot, ctypes.byref(p)))
        return p.value

    def synchronize(self):
        check(lib().pmf_synchronize(self._h))

    def timing(self, on: bool):
        """Bracket every step-kernel launch with HIP events on the state's stream."""
        check(lib().pmf_timing_enable(self._h, 1 if on else 0))

    def timing_read(self):
        """(summed step-kernel milliseconds, number of timed launches) since timing(True)."""
        ms, n = ctypes.c_double(), ctypes.c_int64()
        check(lib().pmf_timing_read(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


def allocate_state(dims, dx, dtype=np.float32, gamma=5.0 / 3.0, smallr=1e-6, pfl=1e-7, llf_dmin=1e-4,
                   llf_pmin=0.0, use_hlld=True, riemann=None, bc="periodic", bcs=None, recon="plm",
                   mhd_solver="glm", workspace=False) -> MHDState:
    """state.jl:109-141. ``dtype`` is np.float32 (the bulk precision) or np.float64.
    ``workspace=True`` allocates no fields: the state then steps caller-owned arrays
    through ``step_into`` / ``max_wavespeed_of``."""
    if mhd_solver not in ("glm", "ct"):
        raise MHDError(f"unknown mhd_solver :{mhd_solver} (have :glm, :ct)")
    bcs = tuple(bcs) if bcs is not None else (bc, bc, bc)
    if mhd_solver == "ct" and any(b != "periodic" for b in bcs):
        raise MHDError(f"mhd_solver=:ct requires all-periodic boundaries (bcs={bcs}); "
                       "constrained transport is periodic-only")
    rsym = riemann if riemann is not None else ("hlld" if use_hlld else "llf")
    p = pmf_params()
    nx, ny, nz = (int(n) for n in dims)
    check(lib().pmf_params_default(ctypes.byref(p), nx, ny, nz, float(dx)))
    p.gamma, p.smallr, p.pfl = float(gamma), float(smallr), float(pfl)
    p.llf_dmin, p.llf_pmin = float(llf_dmin), float(llf_pmin)
    p.riemann = _code(_RIEMANN, "MHD riemann", rsym, ":hlld, :hll, :llf")
    p.recon = _code(_RECON, "recon", recon, ":plm, :plm_minmod, :ppm")
    for d in range(3):
        p.bc[d] = _code(_BC, "BC", bcs[d], ":periodic, :outflow, :reflecting")
    p.solver = 1 if mhd_solver == "ct" else 0
    dt = np.dtype(dtype)
    if dt not in (np.float32, np.float64):
        raise MHDError("dtype must be float32 or float64")
    p.precision = 1 if dt == np.float64 else 0
    return MHDState(p, dt, rsym, recon, bcs, mhd_solver, workspace=workspace)


def fields_to_host(s: MHDState):
    return s.fields_to_host()


def _integ(integrator):
    return _code(_INTEG, "integrator", integrator, ":ref, :cube")


def step(s: MHDState, dt, ch=None, glm_cr=GLM_CR, integrator="ref"):
    """driver.jl:20-37 (``step!``)."""
    code = _integ(integrator)
    if s.uses_ct:
        check(lib().pmf_step(s.handle, float(dt), 0.0, 0.0, code))
        return s
    if ch is None:
        raise MHDError("step! needs the GLM cleaning speed `ch` (GLM states)")
    check(lib().pmf_step(s.handle, float(dt), float(ch), float(glm_cr), code))
    return s


def step_ref(s: MHDState, dt, ch, decay):
    """integrator_ref.jl:210-220 (``step_ref!``) with an explicit psi decay factor."""
    check(lib().pmf_step_raw(s.handle, float(dt), float(ch), float(decay), 0))
    return s


def step_cube(s: MHDState, dt, ch, decay):
    """integrator_cube.jl:455-463 (``step_cube!``): the fused tiled kernel."""
    if s.uses_ct:
        raise MHDError("step_cube!: CT is implemented on the :ref integrator only "
                       "(mhd_solver=:ct states cannot use the :cube integrators)")
    check(lib().pmf_step_raw(s.handle, float(dt), float(ch), float(decay), 1))
    return s


def _ptrs(arrays):
    if len(arrays) != 9:
        raise MHDError("nine field arrays are required")
    return (ctypes.c_void_p * 9)(*[int(a) for a in arrays])


def step_into(ws: MHDState, U_ptrs, O_ptrs, dt, ch, decay, integrator="ref", ct_ptr=None):
    """The _cube_launch! seam: one step from nine caller device arrays into nine others (no
    swap), asynchronous on the workspace's stream. Pointers are integer device addresses."""
    check(lib().pmf_step_into(ws.handle, _ptrs(U_ptrs), _ptrs(O_ptrs), ct_ptr, float(dt), float(ch), float(decay),
                              _integ(integrator)))


def max_wavespeed_of(ws: MHDState, U_ptrs):
    v = ctypes.c_double()
    check(lib().pmf_max_wavespeed_of(ws.handle, _ptrs(U_ptrs), ctypes.byref(v)))
    return ws.dtype.type(v.value)


def max_wavespeed(s: MHDState):
    v = ctypes.c_double()
    check(lib().pmf_max_wavespeed(s.handle, ctypes.byref(v)))
    return s.dtype.type(v.value)


class DeviceBuffer:
    """A zero-filled device allocation owned by the library (pmf_buffer_*)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        h = ctypes.c_void_p()
        check(lib().pmf_buffer_alloc(self.nbytes, ctypes.byref(h)))
        self._p = h

    @property
    def ptr(self) -> int:
        return self._p.value

    def upload(self, a, offset: int = 0):
        a = np.ascontiguousarray(a)
        if offset + a.nbytes > self.nbytes:
            raise MHDError("upload past the end of the buffer")
        check(lib().pmf_buffer_copy(ctypes.c_void_p(self.ptr + offset), a.ctypes.data_as(ctypes.c_void_p),
                                    a.nbytes, 1))

    def download(self, dtype, shape, offset: int = 0):
        out = np.empty(shape, dtype=dtype)
        if offset + out.nbytes > self.nbytes:
            raise MHDError("download past the end of the buffer")
        check(lib().pmf_buffer_copy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr + offset),
                                    out.nbytes, 2))
        return out

    def __del__(self):
        p = getattr(self, "_p", None)
        if p is not None and p.value:
            try:
                lib().pmf_buffer_free(p)
            except Exception:
                pass
            self._p = None


def correction_buffer(s: MHDState, host=None) -> DeviceBuffer:
    """The radiation correction buffer of a state: 9 slots of ncells in the state's
    precision (psi, dv_x, dv_y, dv_z, rho perturbation or fallback mask, B_x, B_y, B_z and
    E residuals), zero-filled or uploaded from a (9, ncells) host array."""
    b = DeviceBuffer(9 * s.ncells * s.dtype.itemsize)
    if host is not None:
        b.upload(np.asarray(host, dtype=s.dtype).reshape(9, s.ncells))
    return b


def load_psi(s: MHDState, corr: DeviceBuffer):
    """Copy the state's psi into slot 0 of the correction buffer (what the radiation
    driver does before each coupled step, radiation_coupling.jl:1681)."""
    p = s.device_ptr(8)
    check(lib().pmf_buffer_copy(ctypes.c_void_p(corr.ptr), ctypes.c_void_p(p), s.ncells * s.dtype.itemsize, 3))


def _vp(b):
    return None if b is None else ctypes.c_void_p(b.ptr)


def species_to_fractions(s: MHDState, species1: DeviceBuffer, species2: DeviceBuffer = None, maximum_total=1.0):
    """conserved_species_to_fraction(s)! (passive_transport.jl:5-25, 198-229) with the
    state's density."""
    check(lib().pmf_species_to_fractions(s.handle, _vp(species1), _vp(species2), None, float(maximum_total)))


def advect_species(s: MHDState, species1: DeviceBuffer, scratch1: DeviceBuffer, dt, species2: DeviceBuffer = None,
                   scratch2: DeviceBuffer = None, maximum_total=1.0):
    """advect_species_fraction(s)! (passive_transport.jl:231-265)."""
    check(lib().pmf_advect_species(s.handle, _vp(species1), _vp(species2), _vp(scratch1), _vp(scratch2), float(dt),
                                   float(maximum_total)))


def advect_passive_scalar(s: MHDState, scalar: DeviceBuffer, scratch: DeviceBuffer, dt):
    """advect_passive_scalar! (passive_transport.jl:267-289)."""
    check(lib().pmf_advect_passive_scalar(s.handle, _vp(scalar), _vp(scratch), float(dt)))


def step_drag(s: MHDState, dt, ch, decay, drag_impulse, target_velocity=(0.0, 0.0, 0.0), integrator="ref"):
    """step_ref_drag! / step_cube_drag! (integrator_ref.jl:222-236, integrator_cube.jl:478-503)."""
    v0 = (ctypes.c_double * 3)(*[float(v) for v in target_velocity])
    check(lib().pmf_step_drag(s.handle, float(dt), float(ch), float(decay), float(drag_impulse), v0,
                              _integ(integrator)))
    return s


def step_radiation(s: MHDState, dt, ch, decay, drag_impulse, correction: DeviceBuffer, track_density=True,
                   track_magnetic=True, track_energy=True, check_admissibility=False, fallback=False,
                   global_fallback=False, robust_fallback=False, integrator="ref"):
    """step_ref_radiation! / step_cube_radiation! (integrator_ref.jl:238-258,
    integrator_cube.jl:505-540). psi is read from slot 0 of the correction buffer."""
    flags = source_flags(track_density, track_magnetic, track_energy, check_admissibility, fallback,
                         global_fallback, robust_fallback)
    check(lib().pmf_step_radiation(s.handle, float(dt), float(ch), float(decay), float(drag_impulse),
                                   ctypes.c_void_p(correction.ptr), flags, _integ(integrator)))
    return s


def source_flags(track_density=True, track_magnetic=True, track_energy=True, check_admissibility=False,
                 fallback=False, global_fallback=False, robust_fallback=False) -> int:
    """The PMF_SRC_* bit set of the radiation step's keyword options."""
    return (1 if track_density else 0) | (2 if track_magnetic else 0) | (4 if track_energy else 0) | \
        (8 if check_admissibility else 0) | (16 if fallback else 0) | (32 if global_fallback else 0) | \
        (64 if robust_fallback else 0)


def step_sources_into(ws: MHDState, U_ptrs, O_ptrs, dt, ch, decay, drag_impulse, target_velocity=(0.0, 0.0, 0.0),
                      correction=None, flags=0, integrator="ref", ct_ptr=None):
    """The _cube_drag_launch! / _cube_radiation_launch! seam (integrator_cube.jl:478-540):
    one drag (correction None) or radiation-coupled step from nine caller device arrays into
    nine others, no swap, asynchronous on the state's stream. correction is a DeviceBuffer
    or an integer device address; flags a source_flags() value."""
    v0 = (ctypes.c_double * 3)(*[float(v) for v in target_velocity])
    corr = None if correction is None else ctypes.c_void_p(int(getattr(correction, "ptr", correction)))
    ct = None if ct_ptr is None else ctypes.c_void_p(int(ct_ptr))
    check(lib().pmf_step_sources_into(ws.handle, _ptrs(U_ptrs), _ptrs(O_ptrs), ct, corr, float(dt), float(ch),
                                      float(decay), float(drag_impulse), v0, int(flags), _integ(integrator)))


def import_mhd_raw(s: MHDState, directory):
    """_import_initial_mhd_state! (runner.jl:3868-3883): the state's nine fields from the
    raw float32 handoff files mhd_U_1.f32 ... mhd_U_9.f32 of a directory."""
    check(lib().pmf_import_mhd_raw(s.handle, os.fsencode(directory)))
    return s


def export_mhd_raw(s: MHDState, directory):
    """The nine fields as the runner's raw float32 handoff files (runner.jl:3814-3825)."""
    os.makedirs(directory, exist_ok=True)
    check(lib().pmf_export_mhd_raw(s.handle, os.fsencode(directory)))
    return s


def set_chunk(s: MHDState, zc: int):
    """Planes per z chunk of the fused kernel (0 = built-in choice); results do not depend on it."""
    check(lib().pmf_set_chunk(s.handle, int(zc)))


def set_ct_kernel(s: MHDState, per_cell: bool):
    """CT steps on the per-cell kernel (True) or, where the grid allows (nx a multiple
    of 16), on the fused kernel that solves every face once (False, the default)."""
    check(lib().pmf_set_ct_kernel(s.handle, 1 if per_cell else 0))


def set_ct_workspace(s: MHDState, fill: bool):
    """Fused CT step: write workspace slots 3-14 and form the edge EMFs in a separate pass,
    as the reference's step does (fill=True), or form the edge EMFs in the step kernel and
    leave slots 3-14 alone (the default); cells, faces and edge EMFs are equal either way."""
    check(lib().pmf_set_ct_workspace(s.handle, 1 if fill else 0))


def autotune(s: MHDState, reps: int = 5):
    """Time the fused kernel over z-chunk depths on this state (scratch is overwritten) and
    keep the fastest; returns (zc, ms per step)."""
    zc, ms = ctypes.c_int32(), ctypes.c_double()
    check(lib().pmf_autotune(s.handle, int(reps), ctypes.byref(zc), ctypes.byref(ms)))
    return zc.value, ms.value


def compute_dt(s: MHDState, cfl=0.4):
    """timestep.jl:43-46 -> (dt, smax) in the state's precision."""
    dt, sm = ctypes.c_double(), ctypes.c_double()
    check(lib().pmf_compute_dt(s.handle, float(cfl), ctypes.byref(dt), ctypes.byref(sm)))
    return s.dtype.type(dt.value), s.dtype.type(sm.value)


def evolve(s: MHDState, tfinal, cfl=0.4, glm_ch_fac=1.0, glm_cr=GLM_CR, integrator="ref", callback=None):
    """driver.jl:48-71 (``evolve!``) -> (t, nsteps)."""
    code = _integ(integrator)
    if callback is None:
        t, n = ctypes.c_double(), ctypes.c_int64()
        check(lib().pmf_evolve(s.handle, float(tfinal), float(cfl), float(glm_ch_fac), float(glm_cr), code,
                               ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value
    if not glm_ch_fac > 0:
        raise MHDError("glm_ch_fac must be positive")
    if not glm_cr >= 0:
        raise MHDError("glm_cr must be non-negative")
    T = s.dtype.type
    t, n, tf = 0.0, 0, float(tfinal)
    while t < tf * (1 - 1e-9):
        dt, smax = compute_dt(s, cfl)
        if s.uses_ct:
            ch, dtf = T(0), float(dt)
        else:
            ch = T(glm_ch_fac) * T(smax)
            dtf = float(dt) * float(smax) / float(max(T(smax), ch))
        if t + dtf > tf:
            dtf = tf - t
        step(s, dtf, ch=ch, glm_cr=glm_cr, integrator=integrator)
        t += dtf
        n += 1
        callback(s, t, n)
    return t, n


def advance(s: MHDState, nsteps: int, cfl=0.4, glm_ch_fac=1.0, glm_cr=GLM_CR):
    """nsteps CFL steps of evolve!'s rule (driver.jl:48-71) on the fused kernel with every
    dt computed on the device (no host round trip per step); equal to nsteps
    compute_dt + step(..., ch = glm_ch_fac * smax) pairs. Returns (elapsed t, last dt)."""
    t, dt = ctypes.c_double(), ctypes.c_double()
    check(lib().pmf_advance(s.handle, int(nsteps), float(cfl), float(glm_ch_fac), float(glm_cr), ctypes.byref(t),
                            ctypes.byref(dt)))
    return t.value, dt.value


def max_divb(s: MHDState) -> float:
    v = ctypes.c_double()
    check(lib().pmf_max_divb(s.handle, ctypes.byref(v)))
    return v.value


def divb_norm(s: MHDState):
    """Float64 (L2, Linf) of the centred div B."""
    a, b = ctypes.c_double(), ctypes.c_double()
    check(lib().pmf_divb_norm_fp64(s.handle, ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


def conserved_totals(s: MHDState) -> dict:
    """diagnostics.jl:44-55 (float64, compensated, on the device)."""
    out = (ctypes.c_double * 8)()
    check(lib().pmf_conserved_totals_fp64(s.handle, out))
    names = ("mass", "momx", "momy", "momz", "energy", "Bx", "By", "Bz")
    return {k: out[i] for i, k in enumerate(names)}


def max_divb_face(s: MHDState) -> float:
    v = ctypes.c_double()
    check(lib().pmf_max_divb_face(s.handle, ctypes.byref(v)))
    return v.value


def _host3(s, arrays):
    return [np.ascontiguousarray(np.asarray(a).reshape(-1), dtype=s.dtype) for a in arrays]


def ct_faces_from_potential(s: MHDState, Ax, Ay, Az):
    """ct.jl:322-345."""
    if not s.uses_ct:
        raise MHDError("ct_faces_from_potential! needs a CT state (mhd_solver=:ct)")
    A = _host3(s, (Ax, Ay, Az))
    for a in A:
        if a.size != s.ncells:
            raise MHDError(f"vector-potential array has {a.size} entries, expected {s.ncells}")
    check(lib().pmf_ct_faces_from_potential(s.handle, *[a.ctypes.data_as(ctypes.c_void_p) for a in A]))
    return s


def ct_install_faces(s: MHDState, bx, by, bz):
    b = _host3(s, (bx, by, bz))
    check(lib().pmf_ct_install_faces(s.handle, *[a.ctypes.data_as(ctypes.c_void_p) for a in b]))
    return s


def ct_faces_from_cells(s: MHDState) -> dict:
    """ct.jl:246-308: divergence-free faces from the cell field (host float64 FFT)."""
    if not s.uses_ct:
        raise MHDError("ct_faces_from_cells! needs a CT state (mhd_solver=:ct)")
    nx, ny, nz = s.dims
    Bold = [s.download(5 + d).astype(np.float64).reshape(nz, ny, nx) for d in range(3)]
    faces = faces_from_cells(*Bold)
    ct_install_faces(s, *faces)
    Bnew = [s.download(5 + d).astype(np.float64) for d in range(3)]
    num = sum(np.sum((Bnew[d] - Bold[d].reshape(-1)) ** 2) for d in range(3))
    den = sum(np.sum(Bold[d] ** 2) for d in range(3))
    enew = sum(np.sum(Bnew[d] ** 2) for d in range(3))
    brms = np.sqrt(enew / s.ncells)
    return dict(cell_b_rel_l2_change=float(np.sqrt(num / den)) if den > 0 else 0.0,
                magnetic_energy_ratio=float(enew / den) if den > 0 else 1.0,
                max_divb_face_dx_over_brms=max_divb_face(s) * float(s.dx) / brms if brms > 0 else 0.0)


def riemann_flux(s: MHDState, L, R, dir: int, ch=None):
    """riemann / riemann_ct (glm_physics.jl:317-348) for n faces of (n, 9) primitive states."""
    L = np.ascontiguousarray(L, dtype=s.dtype)
    R = np.ascontiguousarray(R, dtype=s.dtype)
    if L.shape != R.shape or L.ndim != 2 or L.shape[1] != 9:
        raise MHDError("L and R must be (n, 9) arrays of the same shape")
    F = np.empty_like(L)
    chv = 0.0 if ch is None else float(ch)
    check(lib().pmf_riemann_flux_batch(s.handle, L.ctypes.data_as(ctypes.c_void_p), R.ctypes.data_as(ctypes.c_void_p),
                                       int(dir), chv, L.shape[0], F.ctypes.data_as(ctypes.c_void_p)))
    return F


# --- problem initialisers (problems.jl) -------------------------------------------------
def init_alfven_wave(s: MHDState, amp=1e-3, B0=1.0, p0=0.1):
    s.set_conserved(problems.alfven_wave(s.dims, float(s.dx), float(s.gamma), amp, B0, p0))
    return s


def init_orszag_tang(s: MHDState):
    s.set_conserved(problems.orszag_tang(s.dims, float(s.dx), float(s.gamma)))
    return s


def init_brio_wu(s: MHDState):
    s.set_conserved(problems.brio_wu(s.dims, float(s.gamma)))
    return s


def init_turb_field(s: MHDState):
    s.set_conserved(problems.turb_field(s.dims, float(s.dx), float(s.gamma)))
    return s


def init_field_loop(s: MHDState, radius=0.3, amplitude=1e-3, rho0=1.0, p0=1.0, velocity=(1.0, 1.0, 0.0)):
    """problems.jl:148-180 (for CT states the faces are the curl of an edge A_z)."""
    s.set_conserved(problems.field_loop(s.dims, float(s.dx), float(s.gamma), radius, amplitude, rho0, p0, velocity))
    if s.uses_ct:
        Az = problems.field_loop_Az_edge(s.dims, float(s.dx), radius, amplitude)
        z = np.zeros(s.ncells)
        ct_faces_from_potential(s, z, z, Az)
    return s


def init_config(s: MHDState, name: str, seed: int = 1):
    """Benchmark configuration initial condition (problems.CONFIGS) at the state's N."""
    n = s.dims[0]
    if not (s.dims[0] == s.dims[1] == s.dims[2]):
        raise MHDError("benchmark configurations are cubic")
    s.set_conserved(problems.config_state(name, seed=seed, n=n))
    return s
